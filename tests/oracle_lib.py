"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the compiled
reference subset (oracle/_ref/libharmref.so).  Test infrastructure only."""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int8, c_int64, c_size_t, c_uint32, c_void_p

import numpy as np

from harm_cool_amd.params import HarmParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libharmref.so")
EOSREF_SO = os.path.join(ORACLE_DIR, "_ref", "libeosref.so")

NCOUNTERS = 12
COUNTER_NAMES = ["c2p_calls", "newton_iters", "newton_kconv", "fail_2d", "fail_font1",
                 "fail_font2", "fixup_gamma", "fixup_repairs", "fixup_nostencil",
                 "floor_zones", "gamma_cap", "fatal"]

_dp = POINTER(c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "all"], check=True)


class OrcEos(ctypes.Structure):
    _fields_ = [("nr", c_int), ("nt", c_int), ("rho", _dp), ("T", _dp), ("e", _dp), ("p", _dp), ("gam", c_double)]


def orc_eos(tab: dict, gam: float) -> OrcEos:
    """Oracle view of a host table dict (rho, T, e, p); keeps the arrays alive."""
    arrs = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in tab.items()}
    d = OrcEos(nr=arrs["rho"].size, nt=arrs["T"].size, rho=_ptr(arrs["rho"]), T=_ptr(arrs["T"]), e=_ptr(arrs["e"]),
               p=_ptr(arrs["p"]), gam=gam)
    d._keep = arrs
    return d


class OrcCeos(ctypes.Structure):
    _fields_ = [("X", c_int), ("Y", c_int), ("D", c_int), ("n", c_int), ("vx", _dp), ("vy", _dp), ("values", _dp)] + [
        (k, c_double) for k in ("rho_unit", "u_unit", "p_unit", "gam", "rho_min", "rho_max", "T_min", "T_max")]


def orc_ceos(vx, vy, values, gam: float, rho_unit: float, u_unit: float, p_unit: float, rho_min=None, rho_max=None,
             T_min=None, T_max=None) -> OrcCeos:
    """Oracle view of host caches sharing the grid vx x vy: values [n, X, Y, D]
    (cgs); ranges default to the axis ends.  Keeps the arrays alive."""
    vx, vy = np.ascontiguousarray(vx, dtype=np.float64), np.ascontiguousarray(vy, dtype=np.float64)
    values = np.ascontiguousarray(values, dtype=np.float64)
    n, X, Y, D = values.shape
    d = OrcCeos(X=X, Y=Y, D=D, n=n, vx=_ptr(vx), vy=_ptr(vy), values=_ptr(values), rho_unit=rho_unit, u_unit=u_unit,
                p_unit=p_unit, gam=gam, rho_min=vx[0] if rho_min is None else rho_min,
                rho_max=vx[-1] if rho_max is None else rho_max, T_min=vy[0] if T_min is None else T_min,
                T_max=vy[-1] if T_max is None else T_max)
    d._keep = (vx, vy, values)
    return d


_lib = None


class OrcDiagRec(ctypes.Structure):
    """orc_diag_rec (oracle/harm_oracle.h)."""
    _fields_ = [(n, c_double) for n in (
        "t", "rmed", "pp", "e", "mdot", "edot", "ldot", "lnu", "lnu_rho", "lnu_rho_2", "lnu2",
        "mdot_in", "mdot_out", "mass_start", "mass_curr", "mass_in", "mass_out", "divbmax")] + [
        ("imax", ctypes.c_int32 * 3), ("flux_valid", ctypes.c_int32)]


class OrcRestartHdr(ctypes.Structure):
    """orc_restart_hdr (oracle/harm_oracle.h), restart.c:105-138 in file order."""
    _fields_ = [
        ("N1", ctypes.c_int32), ("N2", ctypes.c_int32), ("N3", ctypes.c_int32), ("nstep", ctypes.c_int32),
        ("t", c_double), ("tf", c_double), ("a", c_double), ("gam", c_double), ("cour", c_double),
        ("DTd", c_double), ("DTl", c_double), ("DTi", c_double),
        ("DTr", ctypes.c_int32), ("dump_cnt", ctypes.c_int32), ("image_cnt", ctypes.c_int32),
        ("rdump_cnt", ctypes.c_int32), ("dt", c_double), ("lim", ctypes.c_int32), ("failed", ctypes.c_int32),
        ("Rin", c_double), ("Rout", c_double), ("hslope", c_double), ("R0", c_double),
        ("mass_start", c_double), ("mass_in", c_double), ("mass_out", c_double),
        ("tdump", c_double), ("timage", c_double), ("tlog", c_double), ("defcon", c_double),
        ("nfailed", ctypes.c_int32), ("ener_size", ctypes.c_int32),
        ("Rbr", c_double), ("cyl_R0", c_double), ("cyl_Th0", c_double),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        _lib = _bind(ctypes.CDLL(ORACLE_SO))
    return _lib


def lib_at(path: str):
    """Another build of the oracle (e.g. one with a planted mistake), bound
    like lib()."""
    return _bind(ctypes.CDLL(path))


def build_mutant(n: int, outdir: str) -> str:
    """The oracle built with ORC_MUTANT=n: the planted mistake n of
    harm_oracle.c (its MUT(n) sites), for tests that show a pin catches it."""
    so = os.path.join(outdir, f"liboracle_mut{n}.so")
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-fopenmp", "-w", "-shared",
                    f"-DORC_MUTANT={n}", "-o", so, os.path.join(ORACLE_DIR, "harm_oracle.c"),
                    os.path.join(ORACLE_DIR, "cache_oracle.c"), "-lm"], check=True)
    return so


def _bind(L):
    if True:
        L.orc_create.restype = c_void_p
        L.orc_create.argtypes = [POINTER(HarmParams)]
        L.orc_destroy.argtypes = [c_void_p]
        L.orc_init_torus.argtypes = [c_void_p]
        L.orc_step.argtypes = [c_void_p]
        L.orc_get_state.argtypes = [c_void_p, _dp, POINTER(c_int8)]
        L.orc_set_state.argtypes = [c_void_p, _dp, POINTER(c_int8)]
        L.orc_get_time.argtypes = [c_void_p, _dp, _dp]
        L.orc_set_time.argtypes = [c_void_p, c_double, c_double]
        L.orc_get_counters.argtypes = [c_void_p, POINTER(c_int64)]
        L.orc_get_grid.argtypes = [c_void_p, _dp, _dp]
        L.orc_get_katm.restype = c_double
        L.orc_get_katm.argtypes = [c_void_p, c_int]
        L.orc_get_geometry.argtypes = [c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.orc_debug_fluxes.argtypes = [c_void_p, c_int, _dp, _dp, _dp, _dp]
        L.orc_debug_fixup_utoprim.argtypes = [c_void_p]
        L.orc_debug_bound.argtypes = [c_void_p]
        L.orc_debug_advance.argtypes = [c_void_p, c_int, _dp]
        L.orc_debug_fixup.argtypes = [c_void_p]
        L.orc_divb_max.restype = c_double
        L.orc_divb_max.argtypes = [c_void_p]
        L.orc_diag.argtypes = [c_void_p, _dp]
        L.orc_diag_log.argtypes = [c_void_p, POINTER(OrcDiagRec)]
        L.orc_set_mass_budget.argtypes = [c_void_p, c_double, c_double, c_double]
        L.orc_tracers_init.restype = ctypes.c_longlong
        L.orc_tracers_init.argtypes = [c_void_p, c_double, c_double]
        for name in ("orc_tracers_dump", "orc_tracers_events"):
            getattr(L, name).restype = ctypes.c_longlong
            getattr(L, name).argtypes = [c_void_p, c_void_p, ctypes.c_longlong]
        L.orc_tracers_get.argtypes = [c_void_p, c_void_p, c_void_p]
        L.orc_tracers_set.argtypes = [c_void_p, c_void_p, c_void_p]
        L.orc_tracers_update.argtypes = [c_void_p, c_double]
        L.orc_set_threads.argtypes = [c_int]
        L.orc_utoprim_2d.argtypes = [_dp, _dp, _dp, c_double, _dp, c_double, POINTER(c_int), POINTER(c_int)]
        L.orc_utoprim_1dvsq2fix1.argtypes = [_dp, _dp, _dp, c_double, _dp, c_double, c_double]
        L.orc_utoprim_1dfix1.argtypes = [_dp, _dp, _dp, c_double, _dp, c_double, c_double]
        L.orc_primtoU_g.argtypes = [_dp, _dp, _dp, c_double, c_double, _dp]
        L.orc_primtoU.argtypes = [_dp, _dp, _dp, c_double, c_double, _dp]
        L.orc_gcov_func.argtypes = [POINTER(HarmParams), _dp, _dp]
        L.orc_dxdxp_func.argtypes = [POINTER(HarmParams), _dp, _dp]
        L.orc_bl_coord_vec.argtypes = [POINTER(HarmParams), _dp, _dp]
        L.orc_conn_func.argtypes = [POINTER(HarmParams), _dp, _dp, _dp]
        L.orc_gdet_func.restype = c_double
        L.orc_gdet_func.argtypes = [_dp]
        L.orc_invert_matrix.argtypes = [_dp, _dp]
        L.orc_slope_lim.restype = c_double
        L.orc_slope_lim.argtypes = [c_int, c_double, c_double, c_double]
        L.orc_kiss_uniform.restype = c_double
        L.orc_kiss_uniform.argtypes = [c_uint32] * 4
        L.orc_face_flux.argtypes = [_dp, _dp, _dp, _dp, c_double, c_int, c_double, c_double, c_double, _dp, _dp]
        L.orc_vchar.argtypes = [_dp, _dp, _dp, c_double, c_int, c_double, _dp, _dp]
        L.orc_eos_table_p.argtypes = [POINTER(OrcEos), c_double, c_double, _dp, _dp, _dp, _dp]
        L.orc_utoprim_5d.argtypes = [_dp, _dp, _dp, c_double, _dp, c_double, POINTER(OrcEos), POINTER(c_int)]
        L.orc_primtoU_eos.argtypes = [_dp, _dp, _dp, c_double, c_double, POINTER(OrcEos), _dp]
        L.orc_dudp_5d.argtypes = [_dp, _dp, _dp, c_double, c_double, POINTER(OrcEos), _dp]
        L.orc_set_eos.argtypes = [c_void_p, POINTER(OrcEos), c_int]
        L.orc_set_eos_cache.argtypes = [c_void_p, POINTER(OrcCeos)]
        L.orc_get_lastT.argtypes = [c_void_p, _dp]
        L.orc_set_lastT.argtypes = [c_void_p, _dp]
        L.orc_ceos_pressure.argtypes = [c_void_p, c_int, c_int, c_int, c_double, c_double]
        L.orc_ceos_pressure.restype = c_double
        L.orc_restart_write.argtypes = [ctypes.c_char_p, POINTER(OrcRestartHdr), _dp, _dp]
        L.orc_restart_read.argtypes = [ctypes.c_char_p, POINTER(OrcRestartHdr), _dp, _dp, c_int]
        L.orc_dump_write.argtypes = [c_void_p, ctypes.c_char_p, POINTER(OrcRestartHdr)]
        L.orc_restart_write_tr.argtypes = [ctypes.c_char_p, POINTER(OrcRestartHdr), _dp, _dp, c_void_p]
    return L


_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_SO)
        R.ref_set_params.argtypes = [c_double] * 5
        R.ref_utoprim_2d.argtypes = [_dp, _dp, _dp, c_double, _dp]
        R.ref_utoprim_1dvsq2fix1.argtypes = [_dp, _dp, _dp, c_double, _dp, c_double]
        R.ref_utoprim_1dfix1.argtypes = [_dp, _dp, _dp, c_double, _dp, c_double]
        R.ref_gcov_func.argtypes = [_dp, _dp]
        R.ref_dxdxp_func.argtypes = [_dp, _dp]
        R.ref_bl_coord_vec.argtypes = [_dp, _dp]
        R.ref_conn_func.argtypes = [_dp, _dp, _dp]
        R.ref_gdet_func.restype = c_double
        R.ref_gdet_func.argtypes = [_dp]
        R.ref_invert_matrix.argtypes = [_dp, _dp]
        R.ref_slope_lim.restype = c_double
        R.ref_slope_lim.argtypes = [c_int, c_double, c_double, c_double]
        R.ref_kiss_uniform.restype = c_double
        R.ref_kiss_uniform.argtypes = [c_uint32] * 4
        R.ref_coord.argtypes = [c_int, c_int, c_int, c_int, _dp, _dp, _dp]
        _ref = R
    return _ref


class Oracle:
    """One oracle simulation context (the restated step_ch state)."""

    def __init__(self, params: HarmParams, L=None):
        self.params = params
        self.L = L if L is not None else lib()
        self.h = self.L.orc_create(ctypes.byref(params))
        if not self.h:
            raise RuntimeError("orc_create failed")
        self.n1, self.n2, self.n3 = params.n1, params.n2, params.n3

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    @property
    def shape(self):
        return (8, self.n1 + 4, self.n2 + 4, self.n3)

    def init_torus(self) -> int:
        return self.L.orc_init_torus(self.h)

    def step(self) -> int:
        return self.L.orc_step(self.h)

    def prims(self) -> np.ndarray:
        a = np.empty(self.shape, dtype=np.float64)
        self.L.orc_get_state(self.h, _ptr(a), None)
        return a

    def pflag(self) -> np.ndarray:
        f = np.empty(self.shape[1:], dtype=np.int8)
        self.L.orc_get_state(self.h, None, f.ctypes.data_as(POINTER(c_int8)))
        return f

    def set_prims(self, a: np.ndarray, pflag: np.ndarray | None = None) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape
        fp = None
        if pflag is not None:
            pflag = np.ascontiguousarray(pflag, dtype=np.int8)
            fp = pflag.ctypes.data_as(POINTER(c_int8))
        self.L.orc_set_state(self.h, _ptr(a), fp)

    def time(self):
        t, dt = c_double(), c_double()
        self.L.orc_get_time(self.h, ctypes.byref(t), ctypes.byref(dt))
        return t.value, dt.value

    def set_time(self, t: float, dt: float) -> None:
        self.L.orc_set_time(self.h, t, dt)

    def counters(self) -> dict:
        c = (c_int64 * NCOUNTERS)()
        self.L.orc_get_counters(self.h, c)
        return dict(zip(COUNTER_NAMES, list(c)))

    def grid(self):
        sx = np.zeros(4)
        dx = np.zeros(4)
        self.L.orc_get_grid(self.h, _ptr(sx), _ptr(dx))
        return sx, dx

    def katm(self) -> np.ndarray:
        return np.array([self.L.orc_get_katm(self.h, i) for i in range(self.n1 + 4)])

    def geometry(self):
        sn1, sn2 = self.n1 + 4, self.n2 + 4
        gcov = np.zeros((sn1, sn2, 5, 4, 4))
        gcon = np.zeros((sn1, sn2, 5, 4, 4))
        gdet = np.zeros((sn1, sn2, 5))
        conn = np.zeros((sn1, sn2, 4, 4, 4))
        coords = np.zeros((sn1, sn2, 4))
        self.L.orc_get_geometry(self.h, _ptr(gcov), _ptr(gcon), _ptr(gdet), _ptr(conn), _ptr(coords))
        return gcov, gcon, gdet, conn, coords

    def debug_fluxes(self, fullstep: int = 0):
        F1 = np.zeros(self.shape)
        F2 = np.zeros(self.shape)
        F3 = np.zeros(self.shape)
        nd = np.zeros(3)
        self.L.orc_debug_fluxes(self.h, fullstep, _ptr(F1), _ptr(F2), _ptr(F3), _ptr(nd))
        return F1, F2, F3, nd

    def debug_fixup_utoprim(self) -> None:
        self.L.orc_debug_fixup_utoprim(self.h)

    def debug_bound(self) -> None:
        self.L.orc_debug_bound(self.h)

    def debug_advance(self, fullstep: int = 0, ph: np.ndarray | None = None) -> np.ndarray:
        if fullstep:
            ph = np.ascontiguousarray(ph, dtype=np.float64).copy()
            assert ph.shape == self.shape
        else:
            ph = np.zeros(self.shape)
        self.L.orc_debug_advance(self.h, fullstep, _ptr(ph))
        return ph

    def debug_fixup(self) -> None:
        self.L.orc_debug_fixup(self.h)

    def set_eos_cache(self, ceos: "OrcCeos | None") -> None:
        rc = self.L.orc_set_eos_cache(self.h, ctypes.byref(ceos) if ceos is not None else None)
        assert rc == 0, rc
        self._ceos = ceos

    def last_T(self) -> np.ndarray:
        T = np.zeros(self.shape[1:])
        assert self.L.orc_get_lastT(self.h, _ptr(T)) == 0
        return T

    def set_last_T(self, T: np.ndarray) -> None:
        T = np.ascontiguousarray(T, dtype=np.float64)
        assert self.L.orc_set_lastT(self.h, _ptr(T)) == 0

    def ceos_pressure(self, h: int, i: int, j: int, rho0: float, u: float) -> float:
        return self.L.orc_ceos_pressure(self.h, h, i, j, rho0, u)

    def divb_max(self) -> float:
        return self.L.orc_divb_max(self.h)

    def diag(self):
        out = np.zeros(4)
        self.L.orc_diag(self.h, _ptr(out))
        return dict(mass=out[0], energy=out[1], mdot_in=out[2], mdot_out=out[3])

    def diag_log(self) -> dict:
        r = OrcDiagRec()
        self.L.orc_diag_log(self.h, ctypes.byref(r))
        d = {n: getattr(r, n) for n, _ in r._fields_}
        d["imax"] = tuple(r.imax)
        d["flux_valid"] = bool(r.flux_valid)
        return d

    def set_mass_budget(self, mass_start: float, mass_in: float = 0.0, mass_out: float = 0.0) -> None:
        self.L.orc_set_mass_budget(self.h, mass_start, mass_in, mass_out)

    def tracers_init(self, rho_unit: float = 1.0, t_unit: float = 1.0) -> int:
        return self.L.orc_tracers_init(self.h, rho_unit, t_unit)

    def _tracer_recs(self, fn):
        from harm_cool_amd import TRACER_DTYPE
        buf = np.zeros(self.n1 * self.n2 * self.n3, dtype=TRACER_DTYPE)
        m = fn(self.h, buf.ctypes.data, buf.size)
        return buf[:m].copy()

    def tracers_dump(self):
        return self._tracer_recs(self.L.orc_tracers_dump)

    def tracers_events(self):
        return self._tracer_recs(self.L.orc_tracers_events)

    def tracers_set(self, X, state) -> None:
        X = np.ascontiguousarray(X, dtype=np.float64)
        st = np.ascontiguousarray(state, dtype=np.int8)
        self.L.orc_tracers_set(self.h, X.ctypes.data, st.ctypes.data)

    def tracers_update(self, dt: float) -> None:
        self.L.orc_tracers_update(self.h, dt)

    def tracers_get(self):
        X = np.empty((3, self.n3, self.n1, self.n2))
        st = np.empty((self.n3, self.n1, self.n2), dtype=np.int8)
        self.L.orc_tracers_get(self.h, X.ctypes.data, st.ctypes.data)
        return X, st


def utoprim_2d(U, gcov, gcon, gdet, prim, gam):
    """Oracle Utoprim_2d on one zone -> (ret, prim_out, n_iter, k_conv)."""
    L = lib()
    U = np.ascontiguousarray(U, dtype=np.float64)
    gcov = np.ascontiguousarray(gcov, dtype=np.float64)
    gcon = np.ascontiguousarray(gcon, dtype=np.float64)
    p = np.array(prim, dtype=np.float64)
    ni, kc = c_int(), c_int()
    r = L.orc_utoprim_2d(_ptr(U), _ptr(gcov), _ptr(gcon), gdet, _ptr(p), gam, ctypes.byref(ni), ctypes.byref(kc))
    return r, p, ni.value, kc.value


class EosRef:
    """The reference's own Fortran EOS chain (oracle/_ref/libeosref.so, built
    from abundance.f90 and its modules by oracle/Makefile); every argument is
    passed by reference, as Fortran does."""

    def __init__(self):
        self.L = ctypes.CDLL(EOSREF_SO)

    def _call(self, name, ins, nout):
        args = [c_double(float(v)) for v in ins] + [c_double() for _ in range(nout)]
        getattr(self.L, name)(*[ctypes.byref(a) for a in args])
        return tuple(a.value for a in args[len(ins):])

    def dfermi(self, dk, eta, theta):
        """(fd, fdeta, fdtheta), gauss_integral.f90:241-349."""
        return self._call("_QMgauss_integralPdfermi", (dk, eta, theta), 3)

    def eos_npe(self, xkt, etan, etap, etae, nn, np_, nee, nep):
        return self._call("_QMeosPeos_npe", (xkt, etan, etap, etae, nn, np_, nee, nep), 3)

    def eos_rad(self, xkt):
        return self._call("_QMeosPeos_rad", (xkt,), 3)

    def eos_neu(self, xkt, b1, b2):
        return self._call("_QMeosPeos_neu", (xkt, b1, b2), 3)

    def eos_he(self, xkt, nhe):
        return self._call("_QMeosPeos_he", (xkt, nhe), 3)

    def abundance(self, rho, T, hdfac):
        """(e, p, Qnu, tau, n_e-, n_e+) of one table point, abundance.f90:6-243."""
        return self._call("abundance_", (rho, T, hdfac), 6)


def eosref_available() -> bool:
    return os.path.exists(EOSREF_SO)


# ---------------------------------------------------------------------------
# Cooling-table cache: the oracle's restatement (oracle/cache_oracle.c) and the
# reference's own cache4d.c/akima.c (oracle/_ref/libcacheref.so).

CACHEREF_SO = os.path.join(ORACLE_DIR, "_ref", "libcacheref.so")


def cacheref_available() -> bool:
    return os.path.exists(CACHEREF_SO)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def cache_lib():
    L = lib()
    L.orc_akima.argtypes = [c_int, _dp, _dp, c_int, _dp, _dp]
    L.orc_akima_der.argtypes = [c_int, _dp, _dp, c_double, _dp, _dp]
    L.orc_lin.argtypes = [c_int, _dp, _dp, c_double]
    L.orc_lin.restype = c_double
    L.orc_cache_get.argtypes = [c_int, c_int, c_int, c_int, POINTER(c_int), c_int, _dp, _dp, _dp, _dp, c_void_p,
                                c_double, c_double, c_double, c_int, c_int, POINTER(c_int), _dp, _dp, c_size_t]
    L.orc_cache_solve_T.argtypes = [c_int, c_int, c_int, c_int, POINTER(c_int), c_int, _dp, _dp, _dp, _dp,
                                    c_void_p, c_double, c_double, c_int, c_double, c_double, c_double, c_int,
                                    c_double, _dp, _dp]
    return L


class OracleCache:
    """The oracle's getvalues() over host arrays (values [z, x, y, d])."""

    def __init__(self, desc, vx, vy, vz, values, counted):
        self.d = desc
        self.vx, self.vy, self.vz, self.values = _f64(vx), _f64(vy), _f64(vz), _f64(values)
        self.counted = np.ascontiguousarray(counted, dtype=np.uint8)
        self.whole = int(desc.wholecachecalculated or bool(np.all(self.counted == 1)))
        self.types = (c_int * 3)(desc.xtype, desc.ytype, desc.ztype)
        self.L = cache_lib()

    def get(self, x, y, z=0.0, zindex=-1, spline=True, dimmask=None, der=False):
        D = self.d.d
        out = np.full(D, np.nan)
        dv = np.full(D, np.nan) if der else None
        mask = None if dimmask is None else (c_int * D)(*[int(m) for m in dimmask])
        p = lambda a: a.ctypes.data_as(_dp)
        s = self.L.orc_cache_get(self.d.x, self.d.y, self.d.z, D, self.types, self.whole, p(self.vx), p(self.vy),
                                 p(self.vz), p(self.values), self.counted.ctypes.data, float(x), float(y), float(z),
                                 int(zindex), int(spline), mask, p(out), None if dv is None else p(dv), 1)
        return s, out, dv


def _solve_args(o):
    p = lambda a: a.ctypes.data_as(_dp)
    return (o.d.x, o.d.y, o.d.z, o.d.d, o.types, o.whole, p(o.vx), p(o.vy), p(o.vz), p(o.values),
            o.counted.ctypes.data)


def oracle_solve_T(o, rho, e, zindex, T0, lo, hi, ntrial=300, tolf=1.0e-6):
    """The oracle's temperature solve on an OracleCache -> (ret, T, berr)."""
    T, b = c_double(), c_double()
    r = o.L.orc_cache_solve_T(*_solve_args(o), float(rho), float(e), int(zindex), float(T0), float(lo), float(hi),
                              int(ntrial), float(tolf), ctypes.byref(T), ctypes.byref(b))
    return r, T.value, b.value


class CacheRef:
    """The reference's Cache4d and akima.c routines through cache_ref_shim.cpp."""

    def __init__(self):
        L = ctypes.CDLL(CACHEREF_SO)
        L.cref_open.restype = c_void_p
        L.cref_open.argtypes = [ctypes.c_char_p, _dp, _dp, _dp]
        L.cref_axes.argtypes = [c_void_p, ctypes.c_char_p, _dp, _dp, _dp]
        L.cref_get.argtypes = [c_void_p, c_double, c_double, c_double, c_int, _dp, _dp, c_int, POINTER(c_int)]
        L.cref_arrays.argtypes = [c_void_p, c_void_p, _dp, _dp, _dp, _dp, c_void_p]
        L.cref_close.argtypes = [c_void_p]
        L.cref_spline_akima.argtypes = [c_int, POINTER(c_int), c_int, _dp, _dp, c_double, _dp]
        L.cref_spline_akima_der.argtypes = [c_int, POINTER(c_int), c_int, _dp, _dp, c_double, _dp, _dp]
        L.cref_linear.argtypes = [c_int, c_int, _dp, _dp, c_double, _dp]
        L.cref_solve_T.argtypes = [c_void_p, c_double, c_double, c_int, c_double, c_double, c_double, c_int,
                                   c_double, _dp, _dp]
        self.L = L

    def desc_size(self) -> int:
        return self.L.cref_desc_size()

    def open(self, path: str):
        h = self.L.cref_open(os.fsencode(path), None, None, None)
        assert h
        return h

    def close(self, h) -> None:
        self.L.cref_close(h)

    def axes(self, desc, missing_path: str):
        vx, vy, vz = np.empty(desc.x), np.empty(desc.y), np.empty(desc.z)
        p = lambda a: a.ctypes.data_as(_dp)
        self.L.cref_axes(ctypes.addressof(desc), os.fsencode(missing_path), p(vx), p(vy), p(vz))
        return vx, vy, vz

    def arrays(self, h, desc_type):
        d = desc_type()
        self.L.cref_arrays(h, ctypes.addressof(d), None, None, None, None, None)
        vx, vy, vz = np.empty(d.x), np.empty(d.y), np.empty(d.z)
        values = np.empty((d.z, d.x, d.y, d.d))
        counted = np.empty((d.z, d.x, d.y), dtype=np.uint8)
        p = lambda a: a.ctypes.data_as(_dp)
        self.L.cref_arrays(h, ctypes.addressof(d), p(vx), p(vy), p(vz), p(values), counted.ctypes.data)
        return {"desc": d, "vx": vx, "vy": vy, "vz": vz, "values": values, "counted": counted}

    def get(self, h, D, x, y, z=0.0, zindex=-1, spline=True, dimmask=None, der=False):
        out = np.full(D, np.nan)
        dv = np.full(D, np.nan) if der else None
        mask = None if dimmask is None else (c_int * D)(*[int(m) for m in dimmask])
        p = lambda a: a.ctypes.data_as(_dp)
        rc = self.L.cref_get(h, float(x), float(y), float(z), int(spline), p(out), None if dv is None else p(dv),
                             int(zindex), mask)
        return rc, out, dv

    def spline_akima(self, t, y, tval, ndim=1, dimmask=None):
        out = np.full(ndim, np.nan)
        t, y = _f64(t), _f64(y)
        mask = None if dimmask is None else (c_int * ndim)(*dimmask)
        self.L.cref_spline_akima(ndim, mask, t.size, t.ctypes.data_as(_dp), y.ctypes.data_as(_dp), float(tval),
                                 out.ctypes.data_as(_dp))
        return out

    def spline_akima_der(self, t, y, tval, ndim=1):
        out, der = np.full(ndim, np.nan), np.full(ndim, np.nan)
        t, y = _f64(t), _f64(y)
        self.L.cref_spline_akima_der(ndim, None, t.size, t.ctypes.data_as(_dp), y.ctypes.data_as(_dp), float(tval),
                                     out.ctypes.data_as(_dp), der.ctypes.data_as(_dp))
        return out, der

    def linear(self, t, y, tval, ndim=1):
        out = np.full(ndim, np.nan)
        t, y = _f64(t), _f64(y)
        self.L.cref_linear(ndim, t.size, t.ctypes.data_as(_dp), y.ctypes.data_as(_dp), float(tval),
                           out.ctypes.data_as(_dp))
        return out

    def solve_T(self, h, rho, e, zindex, T0, lo, hi, ntrial=300, tolf=1.0e-6):
        """The reference's own mnewtc<1> over its getvalues() -> (ret, T, berr)."""
        T, b = c_double(), c_double()
        r = self.L.cref_solve_T(h, float(rho), float(e), int(zindex), float(T0), float(lo), float(hi), int(ntrial),
                                float(tolf), ctypes.byref(T), ctypes.byref(b))
        return r, T.value, b.value
