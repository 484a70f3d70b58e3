"""harm_cool_amd -- MI355X (gfx950) implementation of HARM-COOL's per-timestep
GRMHD update behind the reference's step_ch()/Utoprim_2d() interface.

The compute path is libharmcool.so (hand-written HIP kernels, see csrc/); this
module is the thin Python host over its C ABI (include/harm_cool.h).  There is no
CPU fallback: if the library is missing or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os
import sys
from ctypes import POINTER, c_char_p, c_double, c_int, c_int8, c_int64, c_size_t, c_uint8, c_void_p

import numpy as np

from .params import HarmParams, config1, config2, reference_default, torus_params  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("HARMCOOL_LIB", os.path.join(HERE, "libharmcool.so"))
# The same kernels built with FMA contraction (the fast mode): results within
# rounding of the parity build, not bit-identical to it or to the oracle.
LIB_FAST_PATH = os.path.join(HERE, "libharmcool_fast.so")

NCOUNTERS = 12
COUNTER_NAMES = ["c2p_calls", "newton_iters", "newton_kconv", "fail_2d", "fail_font1",
                 "fail_font2", "fixup_gamma", "fixup_repairs", "fixup_nostencil",
                 "floor_zones", "gamma_cap", "fatal"]
KERNEL_NAMES = ["flux1", "flux2", "flux3", "emf", "update", "bound", "fixutoprim", "stepend",
                "c2p_batch", "mlp", "fontfix", "div", "tracers"]
NKERNELS = len(KERNEL_NAMES)

_dp = POINTER(c_double)


class HarmError(RuntimeError):
    """A library call returned nonzero; .code holds the returned code (11: the
    step's dt fell below 1e-9, 14: a fail() event; see harm_step)."""

    def __init__(self, msg: str, code: int = 1):
        super().__init__(msg)
        self.code = code


class U2POpts(ctypes.Structure):
    _fields_ = [("gam", c_double), ("font_fix", ctypes.c_int32), ("guess_mode", ctypes.c_int32),
                ("mlp", c_void_p), ("K", c_double), ("gate_tau", c_double)]


class EosSpec(ctypes.Structure):
    _fields_ = [("rho_unit", c_double), ("lrho_lo", c_double), ("lrho_hi", c_double), ("lT_lo", c_double),
                ("lT_hi", c_double), ("nr", ctypes.c_int32), ("nt", ctypes.c_int32), ("ye", c_double)]


class EosTable(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int32), ("nt", ctypes.c_int32), ("rho", c_void_p), ("T", c_void_p), ("e", c_void_p),
                ("p", c_void_p), ("gam", c_double)]


class CeosUnits(ctypes.Structure):
    """``harm_ceos_units``: the code units and the cache EOS's ranges (cgs);
    zero ranges take the caches' axis ends."""

    _fields_ = [(n, c_double) for n in ("rho_unit", "u_unit", "p_unit", "rho_min", "rho_max", "T_min", "T_max")]


def reference_units() -> dict:
    """RHO_UNIT, U_UNIT, P_UNIT of the reference's COOL_EOS build
    (cooleos.c:80-135: M_BLH = 3 Msun, M_UNIT = 1.5e-5 Msun, USCALE = 1), in
    the reference's order of operations."""
    GNEWT, MSUN, CL = 6.6742867e-8, 1.9892e33, 2.99792458e10
    M_BLH = 3. * MSUN
    L_UNIT = GNEWT * M_BLH / (CL * CL)
    M_UNIT = 1.5e-5 * MSUN
    RHO_UNIT = M_UNIT / (L_UNIT * L_UNIT * L_UNIT)
    U_UNIT = RHO_UNIT * CL * CL * 1.0
    return dict(rho_unit=RHO_UNIT, u_unit=U_UNIT, p_unit=U_UNIT)


class RestartHdr(ctypes.Structure):
    """Binary image of ``harm_restart_hdr``: the restart header in file order
    (restart.c:105-138)."""

    _fields_ = [
        ("N1", ctypes.c_int32), ("N2", ctypes.c_int32), ("N3", ctypes.c_int32), ("nstep", ctypes.c_int32),
        ("t", c_double), ("tf", c_double), ("a", c_double), ("gam", c_double), ("cour", c_double),
        ("DTd", c_double), ("DTl", c_double), ("DTi", c_double),
        ("DTr", ctypes.c_int32), ("dump_cnt", ctypes.c_int32), ("image_cnt", ctypes.c_int32),
        ("rdump_cnt", ctypes.c_int32),
        ("dt", c_double),
        ("lim", ctypes.c_int32), ("failed", ctypes.c_int32),
        ("Rin", c_double), ("Rout", c_double), ("hslope", c_double), ("R0", c_double),
        ("mass_start", c_double), ("mass_in", c_double), ("mass_out", c_double),
        ("tdump", c_double), ("timage", c_double), ("tlog", c_double), ("defcon", c_double),
        ("nfailed", ctypes.c_int32), ("ener_size", ctypes.c_int32),
        ("Rbr", c_double), ("cyl_R0", c_double), ("cyl_Th0", c_double),
    ]

    @classmethod
    def of(cls, d: "dict | None" = None) -> "RestartHdr":
        h = cls()
        h.ener_size = -1
        for k, v in (d or {}).items():
            setattr(h, k, v)
        return h

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class DiagRec(ctypes.Structure):
    """harm_diag_rec: the ener.out record of diag() (diag.c:59-326)."""
    _fields_ = [(n, c_double) for n in (
        "t", "rmed", "pp", "e", "mdot", "edot", "ldot", "lnu", "lnu_rho", "lnu_rho_2", "lnu2",
        "mdot_in", "mdot_out", "mass_start", "mass_curr", "mass_in", "mass_out", "divbmax")] + [
        ("imax", ctypes.c_int32 * 3), ("flux_valid", ctypes.c_int32)]

    def as_dict(self) -> dict:
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["imax"] = tuple(self.imax)
        d["flux_valid"] = bool(self.flux_valid)
        return d


# harm_tracer_rec: one tracer-file line (tracers.c:470-510) and its tracer
TRACER_DTYPE = np.dtype([("t", "f8"), ("rho", "f8"), ("T", "f8"), ("ye", "f8"), ("r", "f8"), ("th", "f8"),
                         ("ph", "f8"), ("h", "i4"), ("i", "i4"), ("j", "i4"), ("kind", "i4")])
TRACER_NONE, TRACER_LIVE, TRACER_IN, TRACER_OUT, TRACER_JET, TRACER_LOST = range(6)
DUMP_NOALLOC_MAX_ZONES = 8192  # HARM_DUMP_NOALLOC_MAX_ZONES
DUMP_NCOL = 46  # columns of an ASCII dump line (HARM_DUMP_NCOL)
_TRACER_DIRS = {TRACER_IN: "in", TRACER_OUT: "out", TRACER_JET: "jet"}


def tracer_name(root: str, h: int, i: int, j: int, n3: int) -> str:
    """tracer_name (tracers.c:100-107)."""
    if n3 > 2:
        return os.path.join(root, "tracer_%03d_%03d_%03d.dat" % (i, j, h))
    return os.path.join(root, "tracer_%03d_%03d.dat" % (i, j))


def restart_peek(path: str) -> dict:
    """Header of a restart file."""
    h = RestartHdr()
    _check(lib().harm_restart_peek(os.fsencode(path), ctypes.byref(h)), "harm_restart_peek")
    return h.as_dict()


def restart_write_host(path: str, hdr: dict, prim: np.ndarray, temp: "np.ndarray | None" = None) -> None:
    """Write a restart file from host arrays in the canonical layout."""
    h = RestartHdr.of(hdr)
    shape = (8, h.N1 + 4, h.N2 + 4, h.N3)
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    if prim.shape != shape:
        raise ValueError(f"expected prim {shape}, got {prim.shape}")
    tp = None
    if temp is not None:
        temp = np.ascontiguousarray(temp, dtype=np.float64)
        if temp.shape != shape[1:]:
            raise ValueError(f"expected temp {shape[1:]}, got {temp.shape}")
        tp = _ptr(temp)
    _check(lib().harm_restart_write_host(os.fsencode(path), ctypes.byref(h), _ptr(prim), tp), "harm_restart_write_host")


def restart_read_host(path: str, with_temp: bool = False):
    """(header dict, prim[8,N1+4,N2+4,N3], temp or None) of a restart file."""
    hd = restart_peek(path)
    shape = (8, hd["N1"] + 4, hd["N2"] + 4, hd["N3"])
    prim = np.empty(shape, dtype=np.float64)
    temp = np.empty(shape[1:], dtype=np.float64) if with_temp else None
    h, has = RestartHdr(), c_int(0)
    _check(lib().harm_restart_read_host(os.fsencode(path), ctypes.byref(h), _ptr(prim),
                                        _ptr(temp) if temp is not None else None, ctypes.byref(has)),
           "harm_restart_read_host")
    return h.as_dict(), prim, (temp if has.value else None)


def params_from_restart(hdr: dict, **kw) -> HarmParams:
    """Torus run parameters carrying a restart header's grid, metric and run
    controls (other fields from ``torus_params`` defaults or ``kw``)."""
    keys = ("a", "gam", "Rin", "Rout", "hslope", "R0", "Rbr", "cour", "lim", "tf", "defcon")
    args = {k: hdr[k] for k in keys}
    args.update(kw)
    return torus_params(hdr["N1"], hdr["N2"], hdr["N3"], **args)


def eos_spec(**kw) -> EosSpec:
    """The tabulated-EOS workload: density scale 1e11 g/cm^3, log10 rho in
    [6, 13] g/cm^3 and log10 T in [9, 12] K on 200 x 200 points, Ye = 0.5."""
    d = dict(rho_unit=1e11, lrho_lo=6., lrho_hi=13., lT_lo=9., lT_hi=12., nr=200, nt=200, ye=0.5)
    d.update(kw)
    return EosSpec(**d)


def eos_table_build(spec: EosSpec) -> dict:
    """Host arrays of the build's EOS table (harm_eos_table_build)."""
    L = lib()
    rho = np.empty(spec.nr)
    T = np.empty(spec.nt)
    e = np.empty(spec.nr * spec.nt)
    p = np.empty(spec.nr * spec.nt)
    _check(L.harm_eos_table_build(ctypes.byref(spec), _ptr(rho), _ptr(T), _ptr(e), _ptr(p)), "harm_eos_table_build")
    return {"rho": rho, "T": T, "e": e.reshape(spec.nr, spec.nt), "p": p.reshape(spec.nr, spec.nt)}


def c3_hdfac(a: float = 0.9375, r: float = 12.0, m_bh_sun: float = 3.0) -> float:
    """The reference's cooling-table disk-height factor at radius r (in GM/c^2):
    HdFac = 1/Omega = (a + r^1.5) G M / c^3 seconds (cooleos.c:948-962, with its
    constants GNEWT, MSUN, CL of cooleos.c:83-87)."""
    G, MSUN, CL = 6.6742867e-8, 1.9892e33, 2.99792458e10
    return (a + r ** 1.5) * G * m_bh_sun * MSUN / CL ** 3


def fermi_eos_table(spec: EosSpec, hdfac: "float | None" = None, raw: bool = False):
    """Host arrays of the step's tabulated EOS filled by the cooling-table
    generator (harm_eos_table_fermi; abundance() per grid point on the GPU).
    hdfac defaults to c3_hdfac(): the torus's pressure maximum, r = 12.  With
    raw, also the generator's [6][nr*nt] output (e, p, Qnu, tau, n_e-, n_e+)."""
    L = lib()
    hd = c3_hdfac() if hdfac is None else hdfac
    rho = np.empty(spec.nr)
    T = np.empty(spec.nt)
    e = np.empty(spec.nr * spec.nt)
    p = np.empty(spec.nr * spec.nt)
    r = np.empty(6 * spec.nr * spec.nt) if raw else None
    _check(L.harm_eos_table_fermi(ctypes.byref(spec), hd, _ptr(rho), _ptr(T), _ptr(e), _ptr(p),
                                  _ptr(r) if raw else None), "harm_eos_table_fermi")
    tab = {"rho": rho, "T": T, "e": e.reshape(spec.nr, spec.nt), "p": p.reshape(spec.nr, spec.nt)}
    return (tab, r.reshape(6, spec.nr * spec.nt)) if raw else tab


class DeviceEos:
    """A table resident on the device (torch tensors) and its C descriptor."""

    def __init__(self, host: dict, gam: float, device: str = "cuda:0"):
        import torch

        self.t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in host.items()}
        torch.cuda.synchronize()
        self.desc = EosTable(nr=host["rho"].size, nt=host["T"].size, rho=self.t["rho"].data_ptr(),
                             T=self.t["T"].data_ptr(), e=self.t["e"].data_ptr(), p=self.t["p"].data_ptr(), gam=gam)


class CacheDesc(ctypes.Structure):
    """Binary image of ``harm_cache_desc``: the cooling-table cache header, the
    reference's packed 229-byte record (cache4d.h:17-53)."""

    _pack_ = 1
    _fields_ = [
        ("mode", ctypes.c_int32), ("ver", ctypes.c_char * 13),
        ("x", ctypes.c_int32), ("xmin", c_double), ("xmax", c_double), ("xrat", c_double), ("xtype", ctypes.c_int32),
        ("y", ctypes.c_int32), ("ymin", c_double), ("ymax", c_double), ("yrat", c_double), ("ytype", ctypes.c_int32),
        ("z", ctypes.c_int32), ("zmin", c_double), ("zmax", c_double), ("zrat", c_double), ("ztype", ctypes.c_int32),
        ("d", ctypes.c_int32), ("unused", ctypes.c_int32), ("autosave", ctypes.c_int32),
        ("wholecachecalculated", ctypes.c_int32),
        ("xcust", ctypes.c_uint64), ("ycust", ctypes.c_uint64), ("zcust", ctypes.c_uint64),
        ("reserved", ctypes.c_char * 76),
    ]


AXIS_LOG, AXIS_LINEAR, AXIS_CUSTOM, AXIS_NONUNIFORM, AXIS_ROLLING = range(5)
INTERP_LIN, INTERP_SPLINE = 0, 1
CACHE_OUT_OF_RANGE, CACHE_OK, CACHE_NEEDS_RECENTRE, CACHE_MISSING, CACHE_BAD_ZINDEX = range(5)


def cache_desc(x: int, xmin: float, xmax: float, y: int, ymin: float, ymax: float, z: int = 1, zmin: float = 1.0,
               zmax: float = 1.0, d: int = 5, xtype: int = AXIS_LOG, ytype: int = AXIS_LOG, ztype: int = AXIS_LINEAR,
               xrat: float = 0.0, yrat: float = 0.0, zrat: float = 0.0, whole: int = 0) -> CacheDesc:
    """A constant-grid cache header of the reference's version 2.0."""
    return CacheDesc(mode=0, ver=b"AGACACHE_2.0", x=x, xmin=xmin, xmax=xmax, xrat=xrat, xtype=xtype, y=y, ymin=ymin,
                     ymax=ymax, yrat=yrat, ytype=ytype, z=z, zmin=zmin, zmax=zmax, zrat=zrat, ztype=ztype, d=d,
                     wholecachecalculated=whole)


def cache_axis(kind: int, n: int, vmin: float, vmax: float, rat: float = 0.0, cust=None) -> np.ndarray:
    """Axis values of a cache axis type (setindices, cache4d.c:329-388)."""
    out = np.empty(n)
    c = None if cust is None else _ptr(np.ascontiguousarray(cust, dtype=np.float64))
    _check(lib().harm_cache_axis(kind, n, vmin, vmax, rat, c, _ptr(out)), "harm_cache_axis")
    return out


def cache_axes(desc: CacheDesc):
    return (cache_axis(desc.xtype, desc.x, desc.xmin, desc.xmax, desc.xrat),
            cache_axis(desc.ytype, desc.y, desc.ymin, desc.ymax, desc.yrat),
            cache_axis(desc.ztype, desc.z, desc.zmin, desc.zmax, desc.zrat))


def cache_peek(path: str) -> CacheDesc:
    d = CacheDesc()
    _check(lib().harm_cache_peek(os.fsencode(path), ctypes.byref(d)), "harm_cache_peek")
    return d


def cache_read(path: str) -> dict:
    """Header, axes, values [z, x, y, d] and computed flags [z, x, y] of a cache file."""
    d = cache_peek(path)
    vx, vy, vz = np.empty(d.x), np.empty(d.y), np.empty(d.z)
    values = np.empty((d.z, d.x, d.y, d.d))
    counted = np.empty((d.z, d.x, d.y), dtype=np.uint8)
    _check(lib().harm_cache_read(os.fsencode(path), ctypes.byref(d), _ptr(vx), _ptr(vy), _ptr(vz), _ptr(values),
                                 counted.ctypes.data), "harm_cache_read")
    return {"desc": d, "vx": vx, "vy": vy, "vz": vz, "values": values, "counted": counted}


def _cache_arrays(vx, vy, vz, values, counted):
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (vx, vy, vz, values)]
    return a + [np.ascontiguousarray(counted, dtype=np.uint8)]


def cache_write(path: str, desc: CacheDesc, vx, vy, vz, values, counted) -> None:
    vx, vy, vz, values, counted = _cache_arrays(vx, vy, vz, values, counted)
    _check(lib().harm_cache_write(os.fsencode(path), ctypes.byref(desc), _ptr(vx), _ptr(vy), _ptr(vz), _ptr(values),
                                  counted.ctypes.data), "harm_cache_write")


class Cache:
    """A cooling-table cache on the current device (harm_cache_*)."""

    def __init__(self, handle):
        if not handle:
            raise HarmError(f"harm_cache: {lib().harm_last_error().decode(errors='replace')}")
        self.h = handle
        self.desc = CacheDesc()
        w = c_int()
        _check(lib().harm_cache_info(self.h, ctypes.byref(self.desc), ctypes.byref(w)), "harm_cache_info")
        self.whole = bool(w.value)

    @classmethod
    def create(cls, desc: CacheDesc, vx, vy, vz, values, counted) -> "Cache":
        vx, vy, vz, values, counted = _cache_arrays(vx, vy, vz, values, counted)
        return cls(lib().harm_cache_create(ctypes.byref(desc), _ptr(vx), _ptr(vy), _ptr(vz), _ptr(values),
                                           counted.ctypes.data))

    @classmethod
    def open(cls, path: str) -> "Cache":
        return cls(lib().harm_cache_open(os.fsencode(path)))

    def lookup(self, x, y, z=None, zindex=None, mode: int = INTERP_SPLINE, dimmask=None, dery: bool = False):
        """getvalues() for n points (device tensors) -> out [d, n], dery [d, n]
        or None, status [n]; entries not written hold NaN."""
        import torch

        n = x.shape[0]
        _need(x, (n,), "f64", "x")
        _need(y, (n,), "f64", "y")
        if z is not None:
            _need(z, (n,), "f64", "z")
        if zindex is not None:
            _need(zindex, (n,), "i32", "zindex")
        dv = self.desc.d
        out = torch.full((dv, n), float("nan"), dtype=torch.float64, device=x.device)
        der = torch.full((dv, n), float("nan"), dtype=torch.float64, device=x.device) if dery else None
        st = torch.full((n,), -1, dtype=torch.int32, device=x.device)
        mask = None if dimmask is None else (c_int * dv)(*[int(m) for m in dimmask])
        _sync_torch()
        _check(lib().harm_cache_lookup(self.h, mode, mask, x.data_ptr(), y.data_ptr(),
                                       None if z is None else z.data_ptr(),
                                       None if zindex is None else zindex.data_ptr(), out.data_ptr(),
                                       None if der is None else der.data_ptr(), st.data_ptr(), n, None),
               "harm_cache_lookup")
        return out, der, st

    def solve_T(self, rho, e, T, zindex=None, Tlo: float = 0.0, Thi: float = 0.0, ntrial: int = 300,
                tolf: float = 1.0e-6):
        """The COOL_EOS temperature solve (harm_cache_solve_T) for device
        tensors; T (the start) is overwritten. Bounds (both 0) default to the
        ends of the cache's y axis moved in by 1e-6, as the reference sets them.
        -> (berr, ret)."""
        import torch

        n = rho.shape[0]
        _need(rho, (n,), "f64", "rho")
        _need(e, (n,), "f64", "e")
        _need(T, (n,), "f64", "T")
        if zindex is not None:
            _need(zindex, (n,), "i32", "zindex")
        berr = torch.empty(n, dtype=torch.float64, device=rho.device)
        ret = torch.empty(n, dtype=torch.int32, device=rho.device)
        _sync_torch()
        _check(lib().harm_cache_solve_T(self.h, rho.data_ptr(), e.data_ptr(),
                                        None if zindex is None else zindex.data_ptr(), T.data_ptr(), Tlo, Thi,
                                        ntrial, tolf, berr.data_ptr(), ret.data_ptr(), n, None), "harm_cache_solve_T")
        return berr, ret

    def close(self) -> None:
        if self.h:
            lib().harm_cache_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def eos_eval(tab: DeviceEos, rho, u):
    """p, dp/drho, dp/du on device tensors (harm_eos_eval)."""
    import torch

    n = rho.shape[0]
    _need(rho, (n,), "f64", "rho")
    _need(u, (n,), "f64", "u")
    out = [torch.empty(n, dtype=torch.float64, device=rho.device) for _ in range(3)]
    _sync_torch()
    _check(lib().harm_eos_eval(ctypes.byref(tab.desc), rho.data_ptr(), u.data_ptr(), out[0].data_ptr(),
                               out[1].data_ptr(), out[2].data_ptr(), n, None), "harm_eos_eval")
    torch.cuda.synchronize()
    return out


def utoprim5d_batch(U, gcov, gcon, gdet, prim, gam: float, tab: "DeviceEos | None" = None):
    """Batched Utoprim_5d on device tensors (layout as utoprim2d_batch); returns
    (ret int8[n], iters uint8[n])."""
    import torch

    n = gdet.shape[0]
    for t, rows, nm in ((U, 8, "U"), (gcov, 16, "gcov"), (gcon, 16, "gcon"), (prim, 8, "prim")):
        _need(t, (rows, n), "f64", nm)
    ret = torch.empty(n, dtype=torch.int8, device=U.device)
    iters = torch.empty(n, dtype=torch.uint8, device=U.device)
    _sync_torch()
    _check(lib().harm_utoprim5d_batch(U.data_ptr(), gcov.data_ptr(), gcon.data_ptr(), gdet.data_ptr(),
                                      prim.data_ptr(), ret.data_ptr(), iters.data_ptr(), n, gam,
                                      ctypes.byref(tab.desc) if tab is not None else None, None),
           "harm_utoprim5d_batch")
    torch.cuda.synchronize()
    return ret, iters


class StressSpec(ctypes.Structure):
    _fields_ = [(n, c_double) for n in ("a", "hslope", "gam", "r_lo", "r_hi", "th_lo", "th_hi", "rho_lo", "rho_hi",
                                        "urho_lo", "urho_hi", "lorentz_lo", "lorentz_hi", "sigma_lo", "sigma_hi",
                                        "guess_pert")] + [("seed", ctypes.c_uint32 * 4), ("guess_kind", ctypes.c_int32),
                                                          ("pad0", ctypes.c_int32)]


def stress_spec(guess_kind: int = 0, seed=(12345, 65435, 34221, 43543), **kw) -> StressSpec:
    """The inversion stress workload: KS/MKS metric a=0.9375 at r log-uniform
    [1.3, 50], theta uniform (0.01, pi-0.01); rho log-uniform [1e-7, 1], u/rho
    log-uniform [1e-3, 1e2], Lorentz factor uniform [1, 50], sigma log-uniform
    [1e-3, 1e2]; guess = truth with a 1% relative perturbation (kind 0) or cold."""
    d = dict(a=0.9375, hslope=0.3, gam=13. / 9., r_lo=1.3, r_hi=50., th_lo=0.01, th_hi=np.pi - 0.01,
             rho_lo=1e-7, rho_hi=1., urho_lo=1e-3, urho_hi=1e2, lorentz_lo=1., lorentz_hi=50., sigma_lo=1e-3,
             sigma_hi=1e2, guess_pert=0.01)
    d.update(kw)
    s = StressSpec(**d)
    for k in range(4):
        s.seed[k] = seed[k]
    s.guess_kind = guess_kind
    return s


_lib = None


def lib() -> ctypes.CDLL:
    """Load libharmcool.so (built by harm_cool_amd/build.py); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HarmError(f"{LIB_PATH} is missing: run harm_cool_amd/build.py (hipcc, gfx950)")
    # torch bundles its own HIP runtime; loading it first makes the library bind
    # to that same runtime (loaded the other way round, the process would hold
    # two runtimes and the second one finds no device).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


_lib_fast = None


def lib_fast() -> ctypes.CDLL:
    """Load libharmcool_fast.so, the FMA-contracted build of the same library
    (same ABI); raise if absent."""
    global _lib_fast
    if _lib_fast is not None:
        return _lib_fast
    if not os.path.exists(LIB_FAST_PATH):
        raise HarmError(f"{LIB_FAST_PATH} is missing: run harm_cool_amd/build.py (hipcc, gfx950)")
    lib()  # torch's runtime first, as for the parity build
    _lib_fast = _bind(ctypes.CDLL(LIB_FAST_PATH))
    return _lib_fast


def _bind(L: ctypes.CDLL) -> ctypes.CDLL:
    L.harm_last_error.restype = c_char_p
    L.harm_create.restype = c_void_p
    L.harm_create.argtypes = [POINTER(HarmParams), c_int]
    L.harm_destroy.argtypes = [c_void_p]
    for name in ("harm_init_torus", "harm_reset_counters", "harm_debug_bound", "harm_debug_fixup_utoprim",
                 "harm_debug_fixup"):
        getattr(L, name).argtypes = [c_void_p]
    L.harm_step.argtypes = [c_void_p, c_int]
    L.harm_debug_advance.argtypes = [c_void_p, c_int, _dp]
    L.harm_get_prims.argtypes = [c_void_p, _dp, POINTER(c_int8)]
    L.harm_set_prims.argtypes = [c_void_p, _dp, POINTER(c_int8)]
    L.harm_get_time.argtypes = [c_void_p, _dp, _dp]
    L.harm_set_time.argtypes = [c_void_p, c_double, c_double]
    L.harm_get_counters.argtypes = [c_void_p, POINTER(c_int64)]
    L.harm_diag.argtypes = [c_void_p, _dp]
    L.harm_diag_log.argtypes = [c_void_p, POINTER(DiagRec)]
    L.harm_ener_line.argtypes = [POINTER(DiagRec), c_char_p, c_size_t]
    L.harm_set_mass_budget.argtypes = [c_void_p, c_double, c_double, c_double]
    L.harm_tracers_init.argtypes = [c_void_p, c_double, c_double, POINTER(c_int64)]
    L.harm_tracers_stop.argtypes = [c_void_p]
    L.harm_tracers_dump.argtypes = [c_void_p, c_void_p, c_int64, POINTER(c_int64)]
    L.harm_tracers_events.argtypes = [c_void_p, c_void_p, c_int64, POINTER(c_int64)]
    L.harm_tracers_get.argtypes = [c_void_p, c_void_p, c_void_p]
    L.harm_tracers_update.argtypes = [c_void_p, c_double]
    L.harm_fermi_rules.argtypes = [POINTER(c_double)]
    L.harm_fermi_batch.argtypes = [c_void_p] * 6 + [c_size_t, c_void_p]
    L.harm_eos_species_batch.argtypes = [c_int, c_void_p, c_void_p, c_size_t, c_void_p]
    L.harm_eos_point_batch.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]
    L.harm_eos_point_batch_thread.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]
    L.harm_dump_write.argtypes = [c_void_p, c_char_p, POINTER(RestartHdr)]
    L.harm_dump_columns.argtypes = [c_void_p, c_size_t, c_size_t, c_void_p]
    L.harm_dump_columns_noalloc.argtypes = [c_void_p, c_size_t, c_size_t, c_void_p]
    L.harm_tracers_set.argtypes = [c_void_p, c_void_p, c_void_p]
    L.harm_device_state.argtypes = [c_void_p, POINTER(c_void_p), POINTER(c_void_p), POINTER(c_void_p)]
    L.harm_set_timing.argtypes = [c_void_p, c_int]
    L.harm_get_timing.argtypes = [c_void_p, _dp, POINTER(c_int64)]
    L.harm_debug_fluxes.argtypes = [c_void_p, c_int, _dp, _dp, _dp, _dp]
    L.harm_get_geometry.argtypes = [c_void_p, _dp, _dp, _dp, _dp]
    L.harm_utoprim2d_batch.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_size_t, POINTER(U2POpts), c_void_p]
    L.harm_debug_pow.argtypes = [_dp, _dp, _dp, c_size_t]
    L.Utoprim_2d.argtypes = [_dp, _dp, _dp, c_double, _dp]
    L.harm_set_gamma.argtypes = [c_double]
    # learned Newton seed
    L.harm_mlp_create.restype = c_void_p
    L.harm_mlp_create.argtypes = [c_int]
    L.harm_mlp_destroy.argtypes = [c_void_p]
    L.harm_mlp_init.argtypes = [c_void_p, ctypes.c_uint64]
    _fp = POINTER(ctypes.c_float)
    L.harm_mlp_set_params.argtypes = [c_void_p, _fp]
    L.harm_mlp_get_params.argtypes = [c_void_p, _fp]
    L.harm_mlp_set_norm.argtypes = [c_void_p, _dp]
    L.harm_mlp_get_norm.argtypes = [c_void_p, _dp]
    L.harm_mlp_set_vsq_from_w.argtypes = [c_void_p, c_int]
    L.harm_mlp_set_loss_weights.argtypes = [c_void_p, ctypes.c_float, ctypes.c_float]
    L.harm_mlp_fit_norm.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t]
    L.harm_mlp_train.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_int, c_int, ctypes.c_float,
                                 ctypes.c_float, ctypes.c_uint64, _fp]
    L.harm_mlp_train_batch.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int, ctypes.c_float]
    L.harm_mlp_grad.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int, _fp, _dp]
    L.harm_mlp_forward.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int]
    L.harm_mlp_seed.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p,
                                c_void_p, c_void_p]
    L.harm_mlp_pairs.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_double,
                                 c_void_p, c_void_p, c_void_p, c_void_p]
    L.harm_stress_batch.argtypes = [POINTER(StressSpec), c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p]
    L.harm_eos_table_build.argtypes = [POINTER(EosSpec), _dp, _dp, _dp, _dp]
    L.Utoprim_5d.argtypes = [_dp, _dp, _dp, c_double, _dp]
    L.harm_set_u2p_eos.argtypes = [c_void_p]
    L.harm_eos_table_fermi.argtypes = [POINTER(EosSpec), c_double, _dp, _dp, _dp, _dp, _dp]
    L.harm_eos_eval.argtypes = [POINTER(EosTable), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t,
                                c_void_p]
    L.harm_utoprim5d_batch.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_size_t, c_double, POINTER(EosTable), c_void_p]
    L.harm_set_eos.argtypes = [c_void_p, POINTER(EosTable), c_int]
    L.harm_set_eos_cache.argtypes = [c_void_p, POINTER(c_void_p), c_int, POINTER(CeosUnits)]
    L.harm_get_last_T.argtypes = [c_void_p, _dp]
    L.harm_set_last_T.argtypes = [c_void_p, _dp]
    L.harm_get_eos_solves.argtypes = [c_void_p, POINTER(c_int64)]
    L.harm_restart_write.argtypes = [c_void_p, c_char_p, POINTER(RestartHdr)]
    L.harm_restart_read.argtypes = [c_void_p, c_char_p, POINTER(RestartHdr)]
    L.harm_restart_peek.argtypes = [c_char_p, POINTER(RestartHdr)]
    L.harm_restart_write_host.argtypes = [c_char_p, POINTER(RestartHdr), _dp, _dp]
    L.harm_restart_read_host.argtypes = [c_char_p, POINTER(RestartHdr), _dp, _dp, POINTER(c_int)]
    L.harm_set_guess.argtypes = [c_void_p, c_int, c_void_p]
    L.harm_set_guess_gate.argtypes = [c_void_p, c_double]
    L.harm_set_capture.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_int]
    L.harm_get_capture_count.argtypes = [c_void_p, POINTER(c_size_t)]
    L.harm_cache_axis.argtypes = [c_int, c_int, c_double, c_double, c_double, _dp, _dp]
    L.harm_cache_peek.argtypes = [c_char_p, POINTER(CacheDesc)]
    L.harm_cache_read.argtypes = [c_char_p, POINTER(CacheDesc), _dp, _dp, _dp, _dp, c_void_p]
    L.harm_cache_write.argtypes = [c_char_p, POINTER(CacheDesc), _dp, _dp, _dp, _dp, c_void_p]
    L.harm_cache_create.argtypes = [POINTER(CacheDesc), _dp, _dp, _dp, _dp, c_void_p]
    L.harm_cache_create.restype = c_void_p
    L.harm_cache_open.argtypes = [c_char_p]
    L.harm_cache_open.restype = c_void_p
    L.harm_cache_info.argtypes = [c_void_p, POINTER(CacheDesc), POINTER(c_int)]
    L.harm_cache_free.argtypes = [c_void_p]
    L.harm_cache_free.restype = None
    L.harm_cache_lookup.argtypes = [c_void_p, c_int, POINTER(c_int), c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]
    L.harm_cache_solve_T.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_int,
                                     c_double, c_void_p, c_void_p, c_size_t, c_void_p]
    return L


def _check(rc: int, what: str, L: "ctypes.CDLL | None" = None) -> None:
    if rc != 0:
        msg = (L or lib()).harm_last_error().decode(errors="replace")
        raise HarmError(f"{what} failed ({rc}): {msg}", rc)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _sync_torch() -> None:
    """The library's streams do not wait for torch's: finish torch's work on
    the tensors handed over before a call reads them."""
    import torch

    torch.cuda.synchronize()


class Harm:
    """One simulation context: the reference's global state (decs.h:310-378)
    resident on the GPU.  Methods mirror the reference driver: init() is the
    torus setup of init.c, step_ch() one predictor/corrector step of step_ch.c."""

    def __init__(self, params: HarmParams, device: int = 0, fast: bool = False):
        self.params = params
        self.fast = fast
        self.L = lib_fast() if fast else lib()
        self.h = self.L.harm_create(ctypes.byref(params), device)
        if not self.h:
            raise HarmError("harm_create failed: " + self.L.harm_last_error().decode(errors="replace"))
        self.n1, self.n2, self.n3 = params.n1, params.n2, params.n3

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.harm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def shape(self):
        return (8, self.n1 + 4, self.n2 + 4, self.n3)

    def init(self) -> None:
        _check(self.L.harm_init_torus(self.h), "harm_init_torus", self.L)

    def step_ch(self, nsteps: int = 1) -> None:
        _check(self.L.harm_step(self.h, nsteps), "harm_step", self.L)

    def prims(self) -> np.ndarray:
        a = np.empty(self.shape, dtype=np.float64)
        _check(self.L.harm_get_prims(self.h, _ptr(a), None), "harm_get_prims", self.L)
        return a

    def pflag(self) -> np.ndarray:
        f = np.empty(self.shape[1:], dtype=np.int8)
        _check(self.L.harm_get_prims(self.h, None, f.ctypes.data_as(POINTER(c_int8))), "harm_get_prims", self.L)
        return f

    def set_prims(self, a: np.ndarray, pflag: np.ndarray | None = None) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"expected {self.shape}, got {a.shape}")
        fp = None
        if pflag is not None:
            pflag = np.ascontiguousarray(pflag, dtype=np.int8)
            fp = pflag.ctypes.data_as(POINTER(c_int8))
        _check(self.L.harm_set_prims(self.h, _ptr(a), fp), "harm_set_prims", self.L)

    def time(self):
        t, dt = c_double(), c_double()
        _check(self.L.harm_get_time(self.h, ctypes.byref(t), ctypes.byref(dt)), "harm_get_time", self.L)
        return t.value, dt.value

    def set_time(self, t: float, dt: float) -> None:
        _check(self.L.harm_set_time(self.h, t, dt), "harm_set_time", self.L)

    def counters(self) -> dict:
        c = (c_int64 * NCOUNTERS)()
        _check(self.L.harm_get_counters(self.h, c), "harm_get_counters", self.L)
        return dict(zip(COUNTER_NAMES, list(c)))

    def reset_counters(self) -> None:
        _check(self.L.harm_reset_counters(self.h), "harm_reset_counters", self.L)

    def diag(self) -> dict:
        out = np.zeros(7)
        _check(self.L.harm_diag(self.h, _ptr(out)), "harm_diag", self.L)
        return dict(mass=out[0], energy=out[1], mdot_in=out[2], mdot_out=out[3],
                    mass_in=out[4], mass_out=out[5], divb_max=out[6])

    def diag_log(self) -> DiagRec:
        """diag(LOG_OUT)'s record (diag.c:59-326): conserved totals, |div B| max
        and its zone, the inner-face fluxes of the last full step, the mass
        budget."""
        r = DiagRec()
        _check(self.L.harm_diag_log(self.h, ctypes.byref(r)), "harm_diag_log", self.L)
        return r

    def ener_line(self, rec: "DiagRec | None" = None) -> str:
        """One ener.out line (diag.c:317-326) of rec, or of a fresh record."""
        r = rec if rec is not None else self.diag_log()
        buf = ctypes.create_string_buffer(1024)
        n = self.L.harm_ener_line(ctypes.byref(r), buf, len(buf))
        if n < 0:
            raise HarmError("harm_ener_line: line longer than its buffer")
        return buf.value.decode()

    def log_out(self, ener_path: str) -> DiagRec:
        """diag(LOG_OUT): append the ener.out line and print the LOG line
        (diag.c:310-327) to stderr."""
        r = self.diag_log()
        with open(ener_path, "a") as f:
            f.write(self.ener_line(r))
        sys.stderr.write("LOG      t=%g \t divbmax: %d %d %d %g\n" % (r.t, r.imax[0], r.imax[1], r.imax[2], r.divbmax))
        return r

    def set_mass_budget(self, mass_start: float, mass_in: float = 0.0, mass_out: float = 0.0) -> None:
        _check(self.L.harm_set_mass_budget(self.h, mass_start, mass_in, mass_out), "harm_set_mass_budget", self.L)

    # -- tracers (tracers.c:109-654) -------------------------------------------
    def tracers_init(self, rho_unit: float = 1.0, t_unit: float = 1.0) -> int:
        """init_tracers(): seed the tracers; every later step advances them."""
        n = c_int64()
        _check(self.L.harm_tracers_init(self.h, rho_unit, t_unit, ctypes.byref(n)), "harm_tracers_init", self.L)
        self._tracer_buf = np.zeros(n.value, dtype=TRACER_DTYPE)
        return n.value

    def tracers_stop(self) -> None:
        _check(self.L.harm_tracers_stop(self.h), "harm_tracers_stop", self.L)

    def _tracer_call(self, fn, what: str) -> np.ndarray:
        buf = getattr(self, "_tracer_buf", None)
        if buf is None:
            raise HarmError(f"{what}: tracers_init has not been called")
        n = c_int64()
        rc = fn(self.h, buf.ctypes.data, buf.size, ctypes.byref(n))
        if rc == -1:
            self._tracer_buf = buf = np.zeros(n.value, dtype=TRACER_DTYPE)
            rc = fn(self.h, buf.ctypes.data, buf.size, ctypes.byref(n))
        _check(rc, what, self.L)
        return buf[:n.value].copy()

    def tracers_dump(self) -> np.ndarray:
        """dump_tracers(): one TRACER_DTYPE record per live tracer, slot order."""
        return self._tracer_call(self.L.harm_tracers_dump, "harm_tracers_dump")

    def tracers_events(self) -> np.ndarray:
        """Final records of the tracers that left since the previous call."""
        return self._tracer_call(self.L.harm_tracers_events, "harm_tracers_events")

    def tracers_update(self, dt: float) -> None:
        """update_tracers(dt) on the current state (harm_step already does it)."""
        _check(self.L.harm_tracers_update(self.h, dt), "harm_tracers_update", self.L)

    def tracers_get(self):
        """(X[3, N3, N1, N2], state[N3, N1, N2]) of every slot."""
        n1, n2, n3 = self.params.n1, self.params.n2, self.params.n3
        X = np.empty((3, n3, n1, n2))
        st = np.empty((n3, n1, n2), dtype=np.int8)
        _check(self.L.harm_tracers_get(self.h, X.ctypes.data, st.ctypes.data), "harm_tracers_get", self.L)
        return X, st

    def tracers_set(self, X: np.ndarray, state: np.ndarray) -> None:
        X = np.ascontiguousarray(X, dtype=np.float64)
        st = np.ascontiguousarray(state, dtype=np.int8)
        n = self.params.n1 * self.params.n2 * self.params.n3
        if X.size != 3 * n or st.size != n:
            raise HarmError("tracers_set: X must hold 3*N1*N2*N3 values and state N1*N2*N3")
        _check(self.L.harm_tracers_set(self.h, X.ctypes.data, st.ctypes.data), "harm_tracers_set", self.L)

    def write_tracer_files(self, recs: np.ndarray, events: "np.ndarray | None" = None,
                           root: str = "tracers") -> None:
        """dump_tracers()/update_tracers() file output (tracers.c:300-380,
        470-510): append each record to its tracer's file; a tracer that left
        gets its last line, then its file moves to root/in, out or jet."""
        n3 = self.params.n3
        for r in recs:
            with open(tracer_name(root, r["h"], r["i"], r["j"], n3), "a") as f:
                f.write(("%32.22e" * 7 + "\n") % tuple(float(r[k]) for k in ("t", "rho", "T", "ye", "r", "th", "ph")))
        for r in events if events is not None else ():
            src = tracer_name(root, r["h"], r["i"], r["j"], n3)
            with open(src, "a") as f:
                f.write(("%32.22e" * 7 + "\n") % tuple(float(r[k]) for k in ("t", "rho", "T", "ye", "r", "th", "ph")))
            d = os.path.join(root, _TRACER_DIRS[int(r["kind"])])
            os.makedirs(d, exist_ok=True)
            os.replace(src, tracer_name(d, r["h"], r["i"], r["j"], n3))

    def set_timing(self, on: bool) -> None:
        _check(self.L.harm_set_timing(self.h, 1 if on else 0), "harm_set_timing", self.L)

    def timing(self) -> dict:
        ms = np.zeros(NKERNELS)
        n = (c_int64 * NKERNELS)()
        _check(self.L.harm_get_timing(self.h, _ptr(ms), n), "harm_get_timing", self.L)
        return {k: (ms[i], n[i]) for i, k in enumerate(KERNEL_NAMES)}

    def debug_fluxes(self, fullstep: int = 0):
        F1 = np.zeros(self.shape)
        F2 = np.zeros(self.shape)
        F3 = np.zeros(self.shape)
        nd = np.zeros(3)
        _check(self.L.harm_debug_fluxes(self.h, fullstep, _ptr(F1), _ptr(F2), _ptr(F3), _ptr(nd)), "harm_debug_fluxes", self.L)
        return F1, F2, F3, nd

    def debug_bound(self) -> None:
        _check(self.L.harm_debug_bound(self.h), "harm_debug_bound", self.L)

    def debug_fixup_utoprim(self) -> None:
        _check(self.L.harm_debug_fixup_utoprim(self.h), "harm_debug_fixup_utoprim", self.L)

    def debug_advance(self, fullstep: int = 0, ph: np.ndarray | None = None) -> np.ndarray:
        """advance() then fixup() of one half of step_ch (step_ch.c:101-103,
        112-114): fullstep 0 returns the predictor's ph (the state unchanged);
        fullstep 1 takes ph and advances the state by dt."""
        if fullstep:
            ph = np.ascontiguousarray(ph, dtype=np.float64)
            assert ph.shape == self.shape
        else:
            ph = np.zeros(self.shape)
        _check(self.L.harm_debug_advance(self.h, fullstep, _ptr(ph)), "harm_debug_advance", self.L)
        return ph

    def debug_fixup(self) -> None:
        _check(self.L.harm_debug_fixup(self.h), "harm_debug_fixup", self.L)

    def restart_write(self, path: str, **driver) -> None:
        """restart_write() (restart.c:73-184); ``driver`` sets the header fields
        the run driver owns (nstep, dump counters, clocks, ener_size, ...)."""
        h = RestartHdr.of(driver)
        _check(self.L.harm_restart_write(self.h, os.fsencode(path), ctypes.byref(h)), "harm_restart_write", self.L)

    def dump_write(self, path: str, **driver) -> None:
        """dump() ASCII file (dump.c:1710-1990); ``driver`` gives nstep, the
        DT* clocks, the dump/image/rdump counters and ``failed``."""
        h = RestartHdr.of(driver)
        _check(self.L.harm_dump_write(self.h, os.fsencode(path), ctypes.byref(h)), "harm_dump_write", self.L)

    def dump_columns(self, z0: int = 0, n: "int | None" = None) -> np.ndarray:
        """The dump's 46 columns of active zones z0..z0+n-1 in (i, j, h) order."""
        total = self.params.n1 * self.params.n2 * self.params.n3
        n = total - z0 if n is None else n
        out = np.empty((n, DUMP_NCOL))
        _check(self.L.harm_dump_columns(self.h, z0, n, out.ctypes.data), "harm_dump_columns", self.L)
        return out

    def dump_columns_noalloc(self, z0: int, n: int) -> np.ndarray:
        """dump_columns staged in the library's fixed device array (no device
        allocation; n <= DUMP_NOALLOC_MAX_ZONES)."""
        out = np.empty((n, DUMP_NCOL))
        _check(self.L.harm_dump_columns_noalloc(self.h, z0, n, out.ctypes.data), "harm_dump_columns_noalloc", self.L)
        return out

    def restart_read(self, path: str) -> dict:
        """restart_init()/restart_read() (restart.c:195-376) into this context;
        returns the header."""
        h = RestartHdr()
        _check(self.L.harm_restart_read(self.h, os.fsencode(path), ctypes.byref(h)), "harm_restart_read", self.L)
        for k in ("cour", "lim", "tf", "defcon"):
            setattr(self.params, k, getattr(h, k))
        return h.as_dict()

    def set_eos(self, tab: "DeviceEos | None", ref_derivs: bool = False) -> None:
        """Tabulated EOS for the step (None: back to the Gamma-law EOS)."""
        _check(self.L.harm_set_eos(self.h, ctypes.byref(tab.desc) if tab is not None else None, int(ref_derivs)),
               "harm_set_eos", self.L)
        self._eos_ref = tab

    def set_eos_cache(self, caches, units: "dict | None" = None) -> None:
        """The reference's COOL_EOS pressure on cooling-table caches for the
        step (harm_set_eos_cache): one Cache for every radius or one per radial
        index (N1+4); units: rho_unit, u_unit, p_unit (default the reference's)
        and rho_min, rho_max, T_min, T_max (default the caches' axis ends).
        None: back to the Gamma-law EOS."""
        if caches is None:
            _check(self.L.harm_set_eos_cache(self.h, None, 0, None), "harm_set_eos_cache", self.L)
            self._ceos_ref = None
            return
        caches = list(caches)
        u = dict(reference_units())
        u.update(units or {})
        cu = CeosUnits(**u)
        arr = (c_void_p * len(caches))(*[c.h for c in caches])
        _check(self.L.harm_set_eos_cache(self.h, arr, len(caches), ctypes.byref(cu)), "harm_set_eos_cache", self.L)
        self._ceos_ref = caches

    def last_T(self) -> np.ndarray:
        """The zones' last temperatures ((N1+4) x (N2+4) x N3; 0 = unset)."""
        T = np.zeros(self.shape[1:])
        _check(self.L.harm_get_last_T(self.h, _ptr(T)), "harm_get_last_T", self.L)
        return T

    def set_last_T(self, T: np.ndarray) -> None:
        T = np.ascontiguousarray(T, dtype=np.float64)
        assert T.shape == self.shape[1:]
        _check(self.L.harm_set_last_T(self.h, _ptr(T)), "harm_set_last_T", self.L)

    def eos_solves(self) -> int:
        n = c_int64()
        _check(self.L.harm_get_eos_solves(self.h, ctypes.byref(n)), "harm_get_eos_solves", self.L)
        return n.value

    def set_guess(self, mode: int, mlp: "Mlp | None" = None, gate: "float | None" = None) -> None:
        """Newton start of the inversion: 0 previous-step primitives (the
        reference's guess), 1 learned seed with previous-step fallback, 2 the
        learned seed only where the first Newton step from the previous-step
        guess exceeds `gate` (|dW/W| + |dv2/v2|)."""
        if gate is not None:
            _check(self.L.harm_set_guess_gate(self.h, gate), "harm_set_guess_gate", self.L)
        _check(self.L.harm_set_guess(self.h, mode, mlp.h if mlp is not None else None), "harm_set_guess", self.L)
        self._mlp_ref = mlp

    def set_capture(self, feat=None, targ=None, valid=None, stride: int = 1) -> None:
        """Capture training pairs into device tensors feat[8,cap] (float64),
        targ[2,cap] (float64), valid[cap] (uint8); feat=None stops capturing."""
        _sync_torch()
        if feat is None:
            _check(self.L.harm_set_capture(self.h, None, None, None, 0, 1), "harm_set_capture", self.L)
            self._cap_ref = None
            return
        cap = feat.shape[1]
        _train_set(feat, targ, valid)
        _check(self.L.harm_set_capture(self.h, feat.data_ptr(), targ.data_ptr(), valid.data_ptr(), cap, stride),
               "harm_set_capture", self.L)
        self._cap_ref = (feat, targ, valid)

    def capture_count(self) -> int:
        n = c_size_t()
        _check(self.L.harm_get_capture_count(self.h, ctypes.byref(n)), "harm_get_capture_count", self.L)
        return n.value

    def geometry(self):
        ncol = (self.n1 + 4) * (self.n2 + 4)
        gcov = np.zeros((5, 16, ncol))
        gcon = np.zeros((5, 16, ncol))
        gdet = np.zeros((5, ncol))
        conn = np.zeros((64, ncol))
        _check(self.L.harm_get_geometry(self.h, _ptr(gcov), _ptr(gcon), _ptr(gdet), _ptr(conn)), "harm_get_geometry", self.L)
        sh = (self.n1 + 4, self.n2 + 4)
        return (gcov.reshape(5, 4, 4, *sh), gcon.reshape(5, 4, 4, *sh), gdet.reshape(5, *sh),
                conn.reshape(4, 4, 4, *sh))


def device_pow(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """The device pow of the Font-fix solvers evaluated on the GPU (test entry)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(x)
    _check(lib().harm_debug_pow(_ptr(x), _ptr(y), _ptr(out), x.size), "harm_debug_pow")
    return out


def device_copy(nbytes: int, repeats: int = 10) -> float:
    """Time a plain device copy of nbytes; returns ms per copy (counter calibration)."""
    L = lib()
    L.harm_debug_copy.argtypes = [c_size_t, c_int, _dp]
    ms = c_double()
    _check(L.harm_debug_copy(nbytes, repeats, ctypes.byref(ms)), "harm_debug_copy")
    return ms.value


def set_gamma(gam: float) -> None:
    lib().harm_set_gamma(gam)


def set_u2p_eos(tab: "DeviceEos | None") -> None:
    """The calling thread's tabulated EOS for Utoprim_5d (None: the Gamma law)."""
    _check(lib().harm_set_u2p_eos(ctypes.byref(tab.desc) if tab is not None else None), "harm_set_u2p_eos")


def _u2p_one(fn, U, gcov, gcon, gdet: float, prim) -> int:
    U = np.ascontiguousarray(U, dtype=np.float64)
    gcov = np.ascontiguousarray(gcov, dtype=np.float64)
    gcon = np.ascontiguousarray(gcon, dtype=np.float64)
    assert prim.dtype == np.float64 and prim.flags.c_contiguous
    return fn(_ptr(U), _ptr(gcov), _ptr(gcon), gdet, _ptr(prim))


def Utoprim_2d(U, gcov, gcon, gdet: float, prim) -> int:
    """The reference's Utoprim_2d (utoprim_2d.c:126) for one zone, run on the GPU.
    prim is updated in place (a float64 numpy array of 8)."""
    return _u2p_one(lib().Utoprim_2d, U, gcov, gcon, gdet, prim)


def Utoprim_5d(U, gcov, gcon, gdet: float, prim) -> int:
    """The reference's Utoprim_5d (utoprim_5d.c:179) for one zone, run on the GPU
    with the thread's EOS (set_gamma, set_u2p_eos).  prim is updated in place."""
    return _u2p_one(lib().Utoprim_5d, U, gcov, gcon, gdet, prim)


MLP_NPARAMS = 9026
MLP_NORM = 20


def _need(t, shape, dtype_name: str, what: str) -> None:
    """The C ABI reads device tensors as dense row-major arrays: reject views."""
    import torch

    dt = {"f64": torch.float64, "f32": torch.float32, "u8": torch.uint8, "i32": torch.int32}[dtype_name]
    if t.dtype != dt or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what}: expected a contiguous {dtype_name} device tensor of shape {tuple(shape)}, "
                         f"got {tuple(t.shape)} {t.dtype} contiguous={t.is_contiguous()}")


def _train_set(feat, targ, valid):
    n = feat.shape[1]
    _need(feat, (8, n), "f64", "feat")
    _need(targ, (2, n), "f64", "targ")
    if valid is not None:
        _need(valid, (n,), "u8", "valid")
    return n


class Mlp:
    """The learned Newton seed network (harm_mlp_* in include/harm_cool.h):
    trained on the GPU by the library's own kernels, run in BF16 on MFMA."""

    def __init__(self, device: int = 0, seed: int = 1):
        self.L = lib()
        self.h = self.L.harm_mlp_create(device)
        if not self.h:
            raise HarmError("harm_mlp_create failed: " + self.L.harm_last_error().decode(errors="replace"))
        if seed != 1:
            self.init(seed)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.harm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self, seed: int) -> None:
        _check(self.L.harm_mlp_init(self.h, seed), "harm_mlp_init")

    def params(self) -> np.ndarray:
        a = np.empty(MLP_NPARAMS, np.float32)
        _check(self.L.harm_mlp_get_params(self.h, a.ctypes.data_as(POINTER(ctypes.c_float))), "harm_mlp_get_params")
        return a

    def set_params(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.size == MLP_NPARAMS
        _check(self.L.harm_mlp_set_params(self.h, a.ctypes.data_as(POINTER(ctypes.c_float))), "harm_mlp_set_params")

    def norm(self) -> np.ndarray:
        a = np.empty(MLP_NORM)
        _check(self.L.harm_mlp_get_norm(self.h, _ptr(a)), "harm_mlp_get_norm")
        return a

    def set_norm(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(self.L.harm_mlp_set_norm(self.h, _ptr(a)), "harm_mlp_set_norm")

    def set_vsq_from_w(self, on: bool) -> None:
        _check(self.L.harm_mlp_set_vsq_from_w(self.h, int(on)), "harm_mlp_set_vsq_from_w")

    def set_loss_weights(self, w_logw: float, w_vsq: float) -> None:
        _check(self.L.harm_mlp_set_loss_weights(self.h, w_logw, w_vsq), "harm_mlp_set_loss_weights")

    def fit_norm(self, feat, targ, valid) -> None:
        _sync_torch()
        n = _train_set(feat, targ, valid)
        _check(self.L.harm_mlp_fit_norm(self.h, feat.data_ptr(), targ.data_ptr(),
                                        valid.data_ptr() if valid is not None else None, n),
               "harm_mlp_fit_norm")

    def train(self, feat, targ, valid, batch: int = 65536, steps: int = 1000, lr0: float = 2e-3,
              lr1: float = 1e-4, seed: int = 7, trace: bool = False):
        _sync_torch()
        n = _train_set(feat, targ, valid)
        lt = np.zeros(max(steps, 1), np.float32) if trace else None
        _check(self.L.harm_mlp_train(self.h, feat.data_ptr(), targ.data_ptr(),
                                     valid.data_ptr() if valid is not None else None, n, batch, steps,
                                     lr0, lr1, seed,
                                     lt.ctypes.data_as(POINTER(ctypes.c_float)) if lt is not None else None),
               "harm_mlp_train")
        return lt

    def train_batch(self, X, T, mask=None, lr: float = 1e-3) -> None:
        _sync_torch()
        n = X.shape[0]
        _need(X, (n, 8), "f32", "X")
        _need(T, (n, 2), "f32", "T")
        if mask is not None:
            _need(mask, (n,), "u8", "mask")
        _check(self.L.harm_mlp_train_batch(self.h, X.data_ptr(), T.data_ptr(),
                                           mask.data_ptr() if mask is not None else None, X.shape[0], lr),
               "harm_mlp_train_batch")

    def grad(self, X, T, mask=None):
        _sync_torch()
        n = X.shape[0]
        _need(X, (n, 8), "f32", "X")
        _need(T, (n, 2), "f32", "T")
        if mask is not None:
            _need(mask, (n,), "u8", "mask")
        g = np.empty(MLP_NPARAMS, np.float32)
        loss = c_double()
        _check(self.L.harm_mlp_grad(self.h, X.data_ptr(), T.data_ptr(), mask.data_ptr() if mask is not None else None,
                                    X.shape[0], g.ctypes.data_as(POINTER(ctypes.c_float)), ctypes.byref(loss)),
               "harm_mlp_grad")
        return loss.value, g

    def forward(self, X, bf16: bool):
        _sync_torch()
        import torch

        _need(X, (X.shape[0], 8), "f32", "X")
        Y = torch.empty((X.shape[0], 2), dtype=torch.float32, device=X.device)
        _check(self.L.harm_mlp_forward(self.h, X.data_ptr(), Y.data_ptr(), X.shape[0], int(bf16)), "harm_mlp_forward")
        return Y

    def seed(self, U, gcov, gcon, gdet, want_seed: bool = True):
        """Raw features [8,n] and, with want_seed, the network's (W0, vsq0)."""
        _sync_torch()
        import torch

        n = gdet.shape[0]
        for t, rows, nm in ((U, 8, "U"), (gcov, 16, "gcov"), (gcon, 16, "gcon")):
            _need(t, (rows, n), "f64", nm)
        _need(gdet, (n,), "f64", "gdet")
        feat = torch.empty((8, n), dtype=torch.float64, device=U.device)
        W0 = torch.empty(n, dtype=torch.float64, device=U.device) if want_seed else None
        v0 = torch.empty(n, dtype=torch.float64, device=U.device) if want_seed else None
        _check(self.L.harm_mlp_seed(self.h, U.data_ptr(), gcov.data_ptr(), gcon.data_ptr(), gdet.data_ptr(), n,
                                    feat.data_ptr(), W0.data_ptr() if want_seed else None,
                                    v0.data_ptr() if want_seed else None, None), "harm_mlp_seed")
        return feat, W0, v0


def utoprim2d_batch(U, gcov, gcon, gdet, prim, gam: float, font_fix: int = 0, K: float = 0.0, stream=None,
                    mlp: Mlp | None = None, gate: float = 0.0):
    """Batched inversion on device tensors (torch, float64, SoA: U[8,n], gcov[16,n],
    gcon[16,n], gdet[n], prim[8,n] in/out).  Returns (ret int8[n], iters uint8[n],
    kconv uint8[n]) as device tensors."""
    import torch

    n = gdet.shape[0]
    for t, rows in ((U, 8), (gcov, 16), (gcon, 16), (prim, 8)):
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != (rows, n):
            raise ValueError("utoprim2d_batch expects contiguous float64 device tensors of shape (rows, n)")
    ret = torch.empty(n, dtype=torch.int8, device=U.device)
    iters = torch.empty(n, dtype=torch.uint8, device=U.device)
    kconv = torch.empty(n, dtype=torch.uint8, device=U.device)
    mode = 0 if mlp is None else (2 if gate > 0 else 1)
    opts = U2POpts(gam=gam, font_fix=font_fix, guess_mode=mode, mlp=mlp.h if mlp is not None else None, K=K,
                   gate_tau=gate)
    s = stream if stream is not None else torch.cuda.current_stream(U.device).cuda_stream
    _check(lib().harm_utoprim2d_batch(U.data_ptr(), gcov.data_ptr(), gcon.data_ptr(), gdet.data_ptr(),
                                      prim.data_ptr(), ret.data_ptr(), iters.data_ptr(), kconv.data_ptr(), n,
                                      ctypes.byref(opts), s), "harm_utoprim2d_batch")
    return ret, iters, kconv


def stress_batch(n: int, spec: StressSpec, device: str = "cuda:0") -> dict:
    """Generate the stress batch on the device (harm_stress_batch)."""
    import torch

    f = lambda *sh: torch.empty(sh, dtype=torch.float64, device=device)
    out = {"U": f(8, n), "gcov": f(16, n), "gcon": f(16, n), "gdet": f(n), "prim_true": f(8, n), "guess": f(8, n)}
    _sync_torch()
    _check(lib().harm_stress_batch(ctypes.byref(spec), n, out["U"].data_ptr(), out["gcov"].data_ptr(),
                                   out["gcon"].data_ptr(), out["gdet"].data_ptr(), out["prim_true"].data_ptr(),
                                   out["guess"].data_ptr(), None), "harm_stress_batch")
    torch.cuda.synchronize()
    return out


def mlp_pairs(U, gcov, gcon, gdet, prim, ret, gam: float):
    """Training pairs (feat[8,n], targ[2,n], valid[n]) from solved zones."""
    import torch

    n = gdet.shape[0]
    for t, rows, nm in ((U, 8, "U"), (gcov, 16, "gcov"), (gcon, 16, "gcon"), (prim, 8, "prim")):
        _need(t, (rows, n), "f64", nm)
    feat = torch.empty((8, n), dtype=torch.float64, device=U.device)
    targ = torch.empty((2, n), dtype=torch.float64, device=U.device)
    valid = torch.empty(n, dtype=torch.uint8, device=U.device)
    _sync_torch()
    _check(lib().harm_mlp_pairs(U.data_ptr(), gcov.data_ptr(), gcon.data_ptr(), gdet.data_ptr(), prim.data_ptr(),
                                ret.data_ptr(), n, gam, feat.data_ptr(), targ.data_ptr(), valid.data_ptr(), None),
           "harm_mlp_pairs")
    return feat, targ, valid
