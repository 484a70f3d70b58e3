"""Restart files written and read by a context on the GPU (restart.c:73-383):
the file equals the oracle's restart_write() of the oracle's state after the same
steps, a context restarted from it continues bit for bit like the original,
parameters the file cannot change are refused, and with a tabulated EOS the
zone temperatures follow the primitives."""
import ctypes
import os

import numpy as np
import pytest

import harm_cool_amd as hc
from harm_cool_amd.params import torus_params
from oracle_lib import Oracle, OrcRestartHdr, lib as olib, orc_eos

pytestmark = pytest.mark.gpu
_dp = ctypes.POINTER(ctypes.c_double)


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


DRIVER = dict(nstep=3, DTd=5.0, DTl=0.5, DTi=1.0, DTr=100, dump_cnt=1, image_cnt=2, rdump_cnt=0, failed=0,
              tdump=5.0, timage=1.0, tlog=0.5, nfailed=0, cyl_R0=0.0, cyl_Th0=0.0, ener_size=4096)


def test_restart_file_matches_oracle_and_resumes_bitwise(tmp_path):
    _torch()
    P = torus_params(32, 32, 64)
    g = hc.Harm(P)
    o = Oracle(P)
    olib().orc_set_threads(8)
    g.init()
    o.init_torus()
    for _ in range(3):
        g.step_ch(1)
        assert o.step() == 0
    path = str(tmp_path / "rdump0")
    g.restart_write(path, **DRIVER)
    h = hc.restart_peek(path)
    assert (h["N1"], h["N2"], h["N3"]) == (32, 32, 64)
    assert (h["t"], h["dt"]) == g.time()
    for k, v in DRIVER.items():
        if k != "ener_size":  # follows the body: the header peek does not read it
            assert h[k] == v, k
    assert h["mass_start"] == g.diag_log().mass_start > 0.0  # the context's, not the driver's
    # the oracle's restart_write() of its own state under the same header
    oh = OrcRestartHdr()
    for k, v in h.items():
        setattr(oh, k, v)
    oh.ener_size = DRIVER["ener_size"]
    po = np.ascontiguousarray(o.prims())
    assert olib().orc_restart_write(os.fsencode(str(tmp_path / "orc")), ctypes.byref(oh),
                                    po.ctypes.data_as(_dp), None) == 0
    a, b = open(path, "rb").read(), open(tmp_path / "orc", "rb").read()
    if a != b:
        k = next((i for i in range(min(len(a), len(b))) if a[i] != b[i]), min(len(a), len(b)))
        pytest.fail(f"files differ at byte {k} of {len(a)}/{len(b)}: {a[k - 40:k + 40]!r} vs {b[k - 40:k + 40]!r}")
    # a new context from the file continues exactly like the original
    g2 = hc.Harm(hc.params_from_restart(h))
    h2 = g2.restart_read(path)
    assert h2["ener_size"] == DRIVER["ener_size"]
    assert np.array_equal(g2.prims(), g.prims()) and g2.time() == g.time()
    for _ in range(2):
        g.step_ch(1)
        g2.step_ch(1)
    assert np.array_equal(g2.prims(), g.prims()) and g2.time() == g.time()
    assert np.array_equal(g2.pflag(), g.pflag())
    g2.close()
    # a context whose metric differs is refused
    g3 = hc.Harm(torus_params(32, 32, 64, a=0.5))
    with pytest.raises(hc.HarmError, match="a differs"):
        g3.restart_read(path)
    g3.close()
    g.close()


def test_restart_with_tabulated_eos_writes_zone_temperatures(tmp_path):
    _torch()
    P = torus_params(16, 16, 64)
    host = hc.eos_table_build(hc.eos_spec(rho_unit=1e8, nr=60, nt=50))
    tab = hc.DeviceEos(host, P.gam)
    g = hc.Harm(P)
    g.set_eos(tab)
    g.init()
    g.step_ch(1)
    path = str(tmp_path / "rdump0")
    g.restart_write(path, ener_size=0)
    h, prim, temp = hc.restart_read_host(path, with_temp=True)
    assert temp is not None
    assert np.array_equal(prim, g.prims())
    e = orc_eos(host, P.gam)
    L = olib()
    rho, u = prim[0].reshape(-1), prim[1].reshape(-1)
    want = np.empty_like(rho)
    for z in range(rho.size):
        p, a, b, T = (ctypes.c_double() for _ in range(4))
        L.orc_eos_table_p(ctypes.byref(e), rho[z], u[z], ctypes.byref(p), ctypes.byref(a), ctypes.byref(b),
                          ctypes.byref(T))
        want[z] = T.value
    assert np.array_equal(temp.reshape(-1), want)
    assert (want > 0).mean() > 0.1
    # the reader skips the temperature block
    g2 = hc.Harm(hc.params_from_restart(h))
    g2.set_eos(tab)
    g2.restart_read(path)
    assert np.array_equal(g2.prims(), g.prims())
    g2.close()
    g.close()


def test_restart_tracer_block(tmp_path, monkeypatch):
    """restart_write_tracers / restart_read_tracers (tracers.c:565-654): the
    block after the ener.out size equals the oracle's, tracer file sizes are
    those of tracers/ under the run directory, and reading it back keeps the
    tracers whose files exist (cut back to the recorded size) at their
    positions."""
    _torch()
    monkeypatch.chdir(tmp_path)
    P = torus_params(16, 16, 64)
    g = hc.Harm(P)
    o = Oracle(P)
    olib().orc_set_threads(8)
    g.init()
    o.init_torus()
    assert g.tracers_init() == o.tracers_init() > 0
    for _ in range(3):
        g.step_ch(1)
        assert o.step() == 0
    recs = g.tracers_dump()
    os.makedirs("tracers")
    g.write_tracer_files(recs[:40], root="tracers")
    g.write_tracer_files(recs[:20], root="tracers")  # the first 20 files hold two lines
    g.restart_write("rdump", ener_size=123, **{k: v for k, v in DRIVER.items() if k != "ener_size"})
    h = hc.restart_peek("rdump")
    oh = OrcRestartHdr()
    for k, v in h.items():
        setattr(oh, k, v)
    oh.ener_size = 123
    po = np.ascontiguousarray(o.prims())
    assert olib().orc_restart_write_tr(b"orc", ctypes.byref(oh), po.ctypes.data_as(_dp), None, o.h) == 0
    a, b = open("rdump", "rb").read(), open("orc", "rb").read()
    assert a == b
    # later lines appended after the checkpoint are cut away on restart
    g.write_tracer_files(recs[:5], root="tracers")
    g2 = hc.Harm(hc.params_from_restart(h))
    g2.tracers_init()
    g2.restart_read("rdump")
    X, st = g.tracers_get()
    X2, st2 = g2.tracers_get()
    keep = np.zeros(st.size, dtype=bool)
    n12 = P.n1 * P.n2
    for r in recs[:40]:
        keep[r["h"] * n12 + r["i"] * P.n2 + r["j"]] = True
    assert np.array_equal(st2.reshape(-1) == 1, keep)
    assert np.array_equal(X2.reshape(3, -1)[:, keep], X.reshape(3, -1)[:, keep])
    for r in recs[:6]:
        lines = open(hc.tracer_name("tracers", r["h"], r["i"], r["j"], P.n3)).read().splitlines()
        assert len(lines) == 2
    g.close()
    g2.close()
