"""The oracle's restatement of dump() (dump.c:1710-1990, the ASCII dump of the
HDFDUMP = 0 build), pinned by the layout and by identities its columns must
satisfy:

  - the header (2D: N2 and the theta origin/spacing; 3D: N3 and the phi ones),
    one fixed-width line per active zone in (i, j, h) order, 46 columns (20
    when the run failed);
  - coordinates and gdet equal the metric tables, primitives read back exactly;
  - u^mu u_mu = -1, b^mu u_mu = 0, vmin <= vmax, c_s^2 = (4/3) p / (rho + 4 p)
    with p = (gam - 1) u, Mach = -u^r / c_s;
  - the dump's own corner |div B| of a single planted zone, including the
    2D interior-corner rule (i < N1, j < N2 in array indices, dump.c:1864) and
    the 3D (i, j-1) B3 term weighted by gdet of column (i, j) (dump.c:1909).
"""
import ctypes
import math
import os

import numpy as np
import pytest

from harm_cool_amd.params import torus_params
from oracle_lib import Oracle, OrcRestartHdr, lib as olib

CENT = 3
W = 32


def _oracle(n1, n2, n3):
    olib().orc_set_threads(8)
    o = Oracle(torus_params(n1, n2, n3))
    assert o.init_torus() == 0
    return o


def read_dump(path):
    """Header tokens and the [zones, columns] body of a dump file."""
    with open(path) as f:
        head = f.readline()
        body = f.read().splitlines()
    ncol = len(body[0]) // W
    assert all(len(ln) == ncol * W for ln in body)
    cols = np.array([[float(ln[k * W:(k + 1) * W]) for k in range(ncol)] for ln in body])
    return head, cols


def _write(o, path, failed=0):
    h = OrcRestartHdr(nstep=7, DTd=5.0, DTl=0.5, DTi=1.0, DTr=100, dump_cnt=3, image_cnt=4, rdump_cnt=1,
                      failed=failed)
    assert olib().orc_dump_write(o.h, os.fsencode(path), ctypes.byref(h)) == 0


@pytest.mark.parametrize("n3", [1, 4])
def test_layout_and_identities(tmp_path, n3):
    o = _oracle(12, 10, n3)
    for _ in range(2):
        assert o.step() == 0
    path = str(tmp_path / "dump000")
    _write(o, path)
    head, c = read_dump(path)
    sx, dx = o.grid()
    P = o.params
    t, dt = o.time()
    expect = ("%32.22e%10d%10d%32.22e%32.22e%32.22e%32.22e%32.22e%10d" % (
        t, 12, n3 if n3 > 1 else 10, sx[1], sx[3] if n3 > 1 else sx[2], dx[1], dx[3] if n3 > 1 else dx[2],
        P.tf, 7))
    assert head.startswith(expect)
    tail = head.rstrip("\n")[-4 * W:]
    rin, rout, hsl, r0 = (float(tail[k * W:(k + 1) * W]) for k in range(4))
    assert (rout, hsl, r0) == (P.Rout, P.hslope, P.R0)
    assert rin == pytest.approx(math.exp(sx[1]) + P.R0, rel=1e-15)  # Rin <= 0 selects the horizon radius
    assert c.shape == (12 * 10 * n3, 46)
    gcov, gcon, gdet, _, coords = o.geometry()
    p = o.prims()
    ii, jj, hh = np.meshgrid(np.arange(2, 14), np.arange(2, 12), np.arange(n3), indexing="ij")
    ii, jj, hh = ii.ravel(), jj.ravel(), hh.ravel()
    assert np.array_equal(c[:, 4:12], p[:, ii, jj, hh].T)
    assert np.array_equal(c[:, 0], sx[1] + ((ii - 2) + 0.5) * dx[1])
    assert np.array_equal(c[:, 2], coords[ii, jj, 1])
    if n3 > 1:
        assert np.array_equal(c[:, 1], sx[3] + (hh + 0.5) * dx[3]) and np.array_equal(c[:, 3], c[:, 1])
    else:
        assert np.array_equal(c[:, 1], sx[2] + ((jj - 2) + 0.5) * dx[2])
        assert np.array_equal(c[:, 3], coords[ii, jj, 2])
    assert np.all(c[:, 12:19] == 0.0)
    assert np.array_equal(c[:, 40], gdet[ii, jj, CENT])
    ucon, ucov, bcon, bcov = c[:, 20:24], c[:, 24:28], c[:, 28:32], c[:, 32:36]
    assert np.abs((ucon * ucov).sum(1) + 1.0).max() <= 1e-12
    bscale = np.abs(bcon * bcov).sum(1) ** 0.5 * np.abs(ucon * ucov).sum(1) ** 0.5 + 1e-300
    assert (np.abs((bcon * ucov).sum(1)) / bscale).max() <= 1e-12
    assert np.all(c[:, 36] <= c[:, 37]) and np.all(c[:, 38] <= c[:, 39])
    gam = P.gam
    pr = (gam - 1.0) * c[:, 5]
    np.testing.assert_array_equal(c[:, 42], c[:, 5] * (gam - 1.0))
    np.testing.assert_allclose(c[:, 43] ** 2, 4.0 / 3.0 * pr / (c[:, 4] + 4.0 * pr), rtol=1e-14)
    np.testing.assert_allclose(c[:, 44], np.sqrt(gcov[ii, jj, CENT, 1, 1]) * ucon[:, 1] / ucon[:, 0], rtol=1e-15)
    np.testing.assert_array_equal(c[:, 45], -c[:, 44] / c[:, 43])
    assert np.all((c[:, 41] >= 0) & (c[:, 41] < 1))
    # a failed run omits the derived block
    _write(o, str(tmp_path / "dumpf"), failed=1)
    _, cf = read_dump(str(tmp_path / "dumpf"))
    assert cf.shape == (c.shape[0], 20) and np.array_equal(cf, c[:, :20])


def _planted(n3, comp, i0, j0, h0, tmp_path):
    o = _oracle(12, 10, n3)
    p = o.prims().copy()
    p[5:8] = 0.0
    p[comp, i0, j0, h0] = 1.0e-3
    o.set_prims(p)
    path = str(tmp_path / f"d{n3}{comp}")
    _write(o, path)
    _, c = read_dump(path)
    _, _, gdet, _, _ = o.geometry()
    _, dx = o.grid()
    divb = c[:, 19].reshape(12, 10, n3)  # active (i, j, h)
    return divb, gdet[:, :, CENT], dx


def test_planted_divb_2d_interior_rule(tmp_path):
    # plant B1 at array (i0, j0); corners (i0..i0+1, j0..j0+1) carry it
    divb, gd, dx = _planted(1, 5, 6, 5, 0, tmp_path)
    v = abs(0.5 * (1.0e-3 * gd[6, 5]) / dx[1])
    nz = {(i + 2, j + 2) for i, j in zip(*np.nonzero(divb[:, :, 0]))}
    assert nz == {(6, 5), (6, 6), (7, 5), (7, 6)}
    assert all(divb[i - 2, j - 2, 0] == v for i, j in nz)
    # a plant in array column 11 reaches corners 11 and 12; only i < N1 = 12 is
    # written, so the corners of column 12 (and 13, the last active one) read 0
    divb, gd, dx = _planted(1, 5, 11, 5, 0, tmp_path)
    nz = {(i + 2, j + 2) for i, j in zip(*np.nonzero(divb[:, :, 0]))}
    assert nz == {(11, 5), (11, 6)}


def test_planted_divb_3d_b3_weights(tmp_path):
    i0, j0, h0 = 6, 5, 1
    divb, gd, dx = _planted(4, 7, i0, j0, h0, tmp_path)
    got = {(h, i + 2, j + 2): divb[i, j, h] for i, j, h in zip(*np.nonzero(divb))}
    base = abs(0.25 * (1.0e-3 * gd[i0, j0]) / dx[3])
    quirk = abs(0.25 * (1.0e-3 * gd[i0, j0 + 1]) / dx[3])
    want = {(h, i, j): base for h in (h0, h0 + 1) for i in (i0, i0 + 1) for j in (j0, j0 + 1)}
    want[(h0, i0, j0 + 1)] = quirk
    assert got.keys() == want.keys()
    for k, v in want.items():
        assert got[k] == v, k
    assert quirk != base
