"""Tracers on the GPU (harm_tracers_*) against the oracle's restatement of
tracers.c:109-560.

Bar: seeding, positions and states bitwise equal after steps of the same
state; removal and dump records equal in kind, slot, t, rho, T and Ye bitwise,
and in r and theta to 1e-15 relative (exp and sin of the device's math library
against glibc's, each within an ulp); removals come back in (step, slot) order.
"""
import numpy as np
import pytest

import harm_cool_amd as hc
from harm_cool_amd.params import torus_params
from oracle_lib import Oracle, lib as olib
from test_tracers_oracle import _velocity_state

pytestmark = pytest.mark.gpu

CASES = {
    "3d_uniform_32x32x64": lambda: torus_params(32, 32, 64),
    "3d_24x16x8": lambda: torus_params(24, 16, 8),
    "2d_64x64": lambda: torus_params(64, 64, 1),
}


def _pair(P):
    olib().orc_set_threads(16)
    return hc.Harm(P), Oracle(P)


def _same_records(a, b, label):
    assert a.size == b.size, label
    for k in ("t", "rho", "T", "ye", "ph", "h", "i", "j", "kind"):
        assert np.array_equal(a[k], b[k]), f"{label}: {k}"
    for k in ("r", "th"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-15, atol=0, err_msg=f"{label}: {k}")


@pytest.mark.parametrize("case", list(CASES))
def test_tracers_follow_the_oracle_through_steps(case):
    g, o = _pair(CASES[case]())
    g.init()
    o.init_torus()
    assert g.tracers_init(2.0, 3.0) == o.tracers_init(2.0, 3.0) > 0
    X0, sg = g.tracers_get()
    Xo, so = o.tracers_get()
    assert np.array_equal(X0, Xo) and np.array_equal(sg, so)
    for s in range(4):
        g.step_ch(1)
        assert o.step() == 0
    same = np.array_equal(g.prims(), o.prims())
    Xg, sg = g.tracers_get()
    Xo, so = o.tracers_get()
    if same:
        assert np.array_equal(Xg, Xo) and np.array_equal(sg, so)
    else:
        assert np.array_equal(sg, so) and np.abs(Xg - Xo).max() <= 1e-12
    assert np.abs(Xg - X0).max() > 0.0  # the tracers moved
    _same_records(g.tracers_dump(), o.tracers_dump(), f"{case} dump")
    _same_records(g.tracers_events(), o.tracers_events(), f"{case} events")
    g.close()


def test_removals_and_affine_advection_match():
    P = torus_params(16, 12, 64)
    g, o = _pair(P)
    g.init()
    o.init_torus()
    g.tracers_init()
    o.tracers_init()
    coef = np.array([[2e-3, 1e-3, 4e-4, 1e-4],
                     [-1e-3, 2e-4, 5e-4, -2e-4],
                     [3e-3, -5e-4, 1e-3, 3e-4]])
    p = _velocity_state(o, coef)
    g.set_prims(p)
    rng = np.random.default_rng(11)
    X, st = o.tracers_get()
    X = X.reshape(3, -1)
    st = st.reshape(-1)
    st[:] = 0
    idx = rng.choice(st.size, 4000, replace=False)
    X[0, idx] = rng.uniform(-1.0, 4.5, idx.size)   # across the whole radial range
    X[1, idx] = rng.uniform(-1.05, 1.05, idx.size)  # past both poles
    X[2, idx] = rng.uniform(-0.1, 2 * np.pi + 0.1, idx.size)
    st[idx] = 1
    o.tracers_set(X, st)
    g.tracers_set(X, st)
    g.set_time(0.5, 0.1)
    o.set_time(0.5, 0.1)
    for dt in (0.3, 30.0):
        g.tracers_update(dt)
        o.tracers_update(dt)
        Xg, sg = g.tracers_get()
        Xo, so = o.tracers_get()
        assert np.array_equal(sg, so) and np.array_equal(Xg, Xo), dt
        _same_records(g.tracers_events(), o.tracers_events(), f"events dt={dt}")
    kinds = np.bincount(sg.reshape(-1), minlength=6)
    assert kinds[2] > 0 and kinds[3] > 0 and kinds[5] > 0, kinds  # IN, OUT and LOST all happened
    _same_records(g.tracers_dump(), o.tracers_dump(), "dump")
    g.close()


def test_tabulated_eos_dump_temperature():
    import ctypes
    from oracle_lib import orc_eos
    P = torus_params(16, 16, 64)
    g, o = _pair(P)
    host = hc.eos_table_build(hc.eos_spec(rho_unit=1e8, nr=120, nt=100))
    dtab = hc.DeviceEos(host, P.gam)
    g.set_eos(dtab)
    oe = orc_eos(host, P.gam)
    olib().orc_set_eos(o.h, ctypes.byref(oe), 0)
    g.init()
    o.init_torus()
    g.tracers_init()
    o.tracers_init()
    g.step_ch(1)
    o.step()
    a, b = g.tracers_dump(), o.tracers_dump()
    _same_records(a, b, "tabulated dump")
    assert np.all(a["ye"] == -1.0) and (a["T"] > 0).mean() > 0.5
    g.close()


def test_tracer_files(tmp_path):
    P = torus_params(16, 16, 64)
    g = hc.Harm(P)
    g.init()
    g.tracers_init()
    g.step_ch(2)
    recs = g.tracers_dump()
    root = str(tmp_path / "tracers")
    import os
    os.makedirs(root)
    g.write_tracer_files(recs[:5], root=root)
    g.write_tracer_files(recs[:5], root=root)
    r = recs[0]
    f = hc.tracer_name(root, r["h"], r["i"], r["j"], P.n3)
    lines = open(f).read().splitlines()
    assert len(lines) == 2 and len(lines[0]) == 7 * 32
    assert float(lines[0][32:64]) == r["rho"]
    ev = recs[1:2].copy()
    ev["kind"] = hc.TRACER_OUT
    g.write_tracer_files(recs[:0], ev, root=root)
    q = ev[0]
    assert os.path.exists(hc.tracer_name(os.path.join(root, "out"), q["h"], q["i"], q["j"], P.n3))
    g.close()
