"""The CPU oracle against golden vectors written by the reference's own compiled
per-zone code (tests/golden/make_golden.py): bitwise equality."""
import ctypes
import os

import numpy as np
import pytest

from harm_cool_amd.params import config1
from oracle_lib import _ptr, lib

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name))


def test_utoprim_2d_matches_reference_vectors():
    g = _load("utoprim.npz")
    L = lib()
    n = len(g["gdet"])
    nbad = 0
    for k in range(n):
        p = g["guess"][k].copy()
        ni, kc = ctypes.c_int(), ctypes.c_int()
        r = L.orc_utoprim_2d(_ptr(g["U"][k]), _ptr(g["gcov"][k]), _ptr(g["gcon"][k]), g["gdet"][k], _ptr(p),
                             float(g["gam"]), ctypes.byref(ni), ctypes.byref(kc))
        if r != g["ret"][k] or not np.array_equal(p, g["prim"][k]):
            nbad += 1
    assert nbad == 0
    # the fixture exercises success, Newton failure (101), u<=0 (5) and U[0]<=0 (-100)
    assert {0, 5, 101, -100} <= set(np.unique(g["ret"]).tolist())


@pytest.mark.parametrize("which", ["font1", "font2"])
def test_font_fix_matches_reference_vectors(which):
    g = _load("utoprim.npz")
    L = lib()
    fn = L.orc_utoprim_1dvsq2fix1 if which == "font1" else L.orc_utoprim_1dfix1
    nbad = 0
    for k in range(len(g["gdet"])):
        p = g["guess"][k].copy()
        r = fn(_ptr(g["U"][k]), _ptr(g["gcov"][k]), _ptr(g["gcon"][k]), g["gdet"][k], _ptr(p), g["K"][k],
               float(g["gam"]))
        if r != g[f"ret_{which}"][k] or not np.array_equal(p, g[f"prim_{which}"][k]):
            nbad += 1
    assert nbad == 0


def test_geometry_matches_reference_vectors():
    g = _load("geometry.npz")
    L = lib()
    P = config1()
    for k in range(len(g["gdet"])):
        X = g["X"][k]
        gc = np.zeros(16)
        L.orc_gcov_func(ctypes.byref(P), _ptr(X), _ptr(gc))
        assert np.array_equal(gc, g["gcov"][k])
        gi = np.zeros(16)
        assert L.orc_invert_matrix(_ptr(gc), _ptr(gi)) == 0
        assert np.array_equal(gi, g["gcon"][k])
        assert L.orc_gdet_func(_ptr(gc)) == g["gdet"][k]
        c = np.zeros(64)
        L.orc_conn_func(ctypes.byref(P), _ptr(X), _ptr(gi), _ptr(c))
        assert np.array_equal(c, g["conn"][k])
        V = np.zeros(4)
        L.orc_bl_coord_vec(ctypes.byref(P), _ptr(X), _ptr(V))
        assert np.array_equal(V, g["bl"][k])


def test_grid_points_match_reference_coord():
    """coord() of the oracle grid reproduces the reference's code coordinates."""
    from oracle_lib import Oracle
    g = _load("geometry.npz")
    o = Oracle(config1())
    sx, dx = o.grid()
    assert np.array_equal(sx, g["startx"]) and np.array_equal(dx, g["dx"])
    gcov, gcon, gdet, conn, coords = o.geometry()
    for (i, j, loc), gc, gd in zip(g["ijloc"], g["gcov"], g["gdet"]):
        assert np.array_equal(gcov[i, j, loc].ravel(), gc)
        assert gdet[i, j, loc] == gd


def test_slope_limiters_and_kiss_match_reference():
    g = _load("interp_kiss.npz")
    L = lib()
    for y, s in zip(g["y"], g["slope"]):
        for lim in range(3):
            assert L.orc_slope_lim(lim, *y) == s[lim]
    for seed, v in zip(g["seeds"], g["kiss"]):
        assert L.orc_kiss_uniform(*[int(x) for x in seed]) == v
