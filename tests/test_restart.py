"""Restart files (restart.c:73-383): the library's writer against the oracle's
restatement of restart_write() byte for byte, exact round trips (23 significant
digits), the reader against the oracle's fscanf reader, layouts the fixed-width
fast path does not take, and malformed files.  Host arrays only (no GPU)."""
import ctypes
import os

import numpy as np
import pytest

import harm_cool_amd as hc
from oracle_lib import OrcRestartHdr, lib as olib

_dp = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


def _hdr(n1=6, n2=8, n3=4, ener=1234, seed=0):
    rng = np.random.default_rng(seed)
    d = {k: float(rng.normal() * 10.0 ** rng.integers(-8, 8)) for k, t in hc.RestartHdr._fields_
         if t is ctypes.c_double}
    d.update(N1=n1, N2=n2, N3=n3, nstep=int(rng.integers(0, 10 ** 6)), DTr=7, dump_cnt=3, image_cnt=11,
             rdump_cnt=2, lim=0, failed=0, nfailed=-5, ener_size=ener)
    return d


def _prims(h, seed=1):
    rng = np.random.default_rng(seed)
    shape = (8, h["N1"] + 4, h["N2"] + 4, h["N3"])
    a = rng.normal(size=shape) * 10.0 ** rng.integers(-300, 300, size=shape)
    flat = a.reshape(-1)
    special = [0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, 1.0, -1.0, 0.1,
               1.0 / 3.0, np.nextafter(1.0, 2.0), 9.999999999999999e22]
    flat[:len(special)] = special
    return a


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _oracle_write(path, h, prim, temp=None):
    oh = OrcRestartHdr()
    for k, v in h.items():
        setattr(oh, k, v)
    rc = olib().orc_restart_write(os.fsencode(str(path)), ctypes.byref(oh), _p(prim),
                                  _p(temp) if temp is not None else None)
    assert rc == 0


def test_header_layouts_agree():
    assert ctypes.sizeof(hc.RestartHdr) == ctypes.sizeof(OrcRestartHdr) == 232
    for (a, _), (b, _) in zip(hc.RestartHdr._fields_, OrcRestartHdr._fields_):
        assert a == b and getattr(hc.RestartHdr, a).offset == getattr(OrcRestartHdr, b).offset


@pytest.mark.parametrize("ener,with_temp", [(1234, False), (-1, False), (0, True), (-1, True)])
def test_writer_matches_oracle_bytes(tmp_path, ener, with_temp):
    h = _hdr(ener=ener)
    prim = _prims(h)
    temp = np.random.default_rng(2).uniform(1e9, 1e12, size=prim.shape[1:]) if with_temp else None
    hc.restart_write_host(str(tmp_path / "lib"), h, prim, temp)
    _oracle_write(tmp_path / "orc", h, prim, temp)
    a = open(tmp_path / "lib", "rb").read()
    b = open(tmp_path / "orc", "rb").read()
    assert a == b
    # header line of 34 fixed-width fields, then 257-byte zone lines
    nz = prim[0].size
    hdr_len = 11 * 10 + 23 * 32 + 1
    assert a[hdr_len - 1:hdr_len] == b"\n"
    tail = (10 if ener >= 0 else 0) + (32 * nz if with_temp else 0)
    assert len(a) == hdr_len + 257 * nz + tail


@pytest.mark.parametrize("with_temp", [False, True])
def test_round_trip_is_exact(tmp_path, with_temp):
    h = _hdr(n1=5, n2=9, n3=6, ener=77 if with_temp else -1, seed=3)
    prim = _prims(h, seed=4)
    temp = np.random.default_rng(5).normal(size=prim.shape[1:]) if with_temp else None
    path = str(tmp_path / "r")
    hc.restart_write_host(path, h, prim, temp)
    h2, p2, t2 = hc.restart_read_host(path, with_temp=True)
    assert np.array_equal(_bits(p2), _bits(prim))  # -0.0 and subnormals included
    for k, v in h.items():
        assert h2[k] == v, k
    if with_temp:
        assert np.array_equal(_bits(t2), _bits(temp))
    else:
        assert t2 is None
    assert hc.restart_peek(path)["nstep"] == h["nstep"]


def test_reader_matches_oracle_reader(tmp_path):
    h = _hdr(ener=99, seed=6)
    prim = _prims(h, seed=7)
    path = str(tmp_path / "r")
    _oracle_write(path, h, prim)
    oh = OrcRestartHdr()
    po = np.empty_like(prim)
    assert olib().orc_restart_read(os.fsencode(path), ctypes.byref(oh), _p(po), None, 1) == 0
    h2, p2, _ = hc.restart_read_host(path)
    assert np.array_equal(_bits(p2), _bits(po))
    assert np.array_equal(_bits(po), _bits(prim))
    assert h2["ener_size"] == oh.ener_size == 99


def test_any_whitespace_layout_reads_token_by_token(tmp_path):
    h = _hdr(n1=4, n2=8, n3=1, ener=5, seed=8)
    prim = _prims(h, seed=9)
    path = str(tmp_path / "r")
    hc.restart_write_host(path, h, prim)
    toks = open(path).read().split()
    # one token per line: the fixed-width body check fails, the token reader runs
    with open(tmp_path / "s", "w") as f:
        f.write("\n".join(toks) + "\n")
    h2, p2, _ = hc.restart_read_host(str(tmp_path / "s"))
    assert np.array_equal(_bits(p2), _bits(prim))
    assert h2["ener_size"] == 5


def test_malformed_files_are_rejected(tmp_path):
    h = _hdr(n1=4, n2=8, n3=2, ener=-1, seed=10)
    prim = _prims(h, seed=11)
    path = str(tmp_path / "r")
    hc.restart_write_host(path, h, prim)
    data = open(path, "rb").read()
    open(tmp_path / "trunc", "wb").write(data[:len(data) - 300])
    with pytest.raises(hc.HarmError, match="truncated"):
        hc.restart_read_host(str(tmp_path / "trunc"))
    open(tmp_path / "hdr", "wb").write(data[:200])
    with pytest.raises(hc.HarmError, match="header"):
        hc.restart_peek(str(tmp_path / "hdr"))
    with pytest.raises(hc.HarmError, match="cannot open"):
        hc.restart_peek(str(tmp_path / "missing"))
    with pytest.raises(ValueError):
        hc.restart_write_host(path, h, prim[:, :-1])


@pytest.mark.parametrize("n", [(60, 60, 16), (124, 124, 66)])
def test_threaded_and_chunked_writer_matches_oracle(tmp_path, n):
    # 65,536 zones spread over all writer threads; 1,081,344 zones span two
    # formatting chunks of 2^20 zones
    h = _hdr(*n, ener=3, seed=12)
    shape = (8, n[0] + 4, n[1] + 4, n[2])
    prim = np.random.default_rng(13).normal(size=shape)
    hc.restart_write_host(str(tmp_path / "lib"), h, prim)
    _oracle_write(tmp_path / "orc", h, prim)
    with open(tmp_path / "lib", "rb") as fa, open(tmp_path / "orc", "rb") as fb:
        while True:
            a, b = fa.read(1 << 24), fb.read(1 << 24)
            assert a == b
            if not a:
                break
    _, p2, _ = hc.restart_read_host(str(tmp_path / "lib"))
    assert np.array_equal(_bits(p2), _bits(prim))
