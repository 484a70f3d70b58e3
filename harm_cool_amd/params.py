"""Run parameters and named presets for the HARM-COOL GRMHD update.

The reference fixes every parameter at compile time (decs.h:59-258, init.c:90-161,
step_ch.c:66-67).  Here they are one runtime struct, ``harm_params`` in
include/harm_cool.h, mirrored by :class:`HarmParams`.  The presets are the
workloads the project measures: the 2D config-1 torus, the 288x256x256 3D torus
and small variants used by the parity tests.
"""
from __future__ import annotations

import ctypes
import math
from ctypes import c_double, c_int32, c_uint32

MC, VANL, MINM = 0, 1, 2  # slope limiters (decs.h:214-216)
PPM = 3  # piecewise-parabolic reconstruction on MC slopes (no counterpart in the reference)


class HarmParams(ctypes.Structure):
    """Binary image of ``harm_params`` (include/harm_cool.h)."""

    _fields_ = [
        ("n1", c_int32), ("n2", c_int32), ("n3", c_int32),
        ("lim", c_int32), ("font_fix", c_int32), ("pad0", c_int32),
        ("a", c_double), ("gam", c_double),
        ("rin", c_double), ("rmax", c_double),
        ("kappa", c_double), ("beta", c_double),
        ("Rin", c_double), ("Rout", c_double),
        ("hslope", c_double), ("R0", c_double), ("Rbr", c_double),
        ("cour", c_double), ("dt0", c_double), ("tf", c_double), ("defcon", c_double),
        ("hllf", c_double), ("laxf", c_double),
        ("pert_amp", c_double),
        ("seed", c_uint32 * 4),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("seed", "pad0")}
        d["seed"] = list(self.seed)
        return d


def torus_params(n1: int, n2: int, n3: int, **kw) -> HarmParams:
    """Fishbone-Moncrief torus in MKS coordinates.

    Defaults follow the measured workloads: a=0.9375, Gamma=13/9, rin=6,
    rmax=12, Rout=40, hslope=0.3, beta=100, kappa=1e-3, MC limiter, HLL flux,
    Courant 0.8, dt0=1e-5 and the reference's u perturbation of
    4e-2*(U[0,1]-0.5) with KISS seeds {12345,65435,34221,43543} (init.c:266-274).
    Rin <= 0 selects 0.98*(1+sqrt(1-a^2)) (init.c:125).
    """
    p = HarmParams()
    p.n1, p.n2, p.n3 = n1, n2, n3
    p.lim = kw.pop("lim", MC)
    p.font_fix = kw.pop("font_fix", 1)
    p.a = kw.pop("a", 0.9375)
    p.gam = kw.pop("gam", 13.0 / 9.0)
    p.rin = kw.pop("rin", 6.0)
    p.rmax = kw.pop("rmax", 12.0)
    p.kappa = kw.pop("kappa", 1.0e-3)
    p.beta = kw.pop("beta", 100.0)
    p.Rin = kw.pop("Rin", 0.0)
    p.Rout = kw.pop("Rout", 40.0)
    p.hslope = kw.pop("hslope", 0.3)
    p.R0 = kw.pop("R0", 0.0)
    p.Rbr = kw.pop("Rbr", -1.0)
    p.cour = kw.pop("cour", 0.8)
    p.dt0 = kw.pop("dt0", 1.0e-5)
    p.tf = kw.pop("tf", 1.0e30)
    p.defcon = kw.pop("defcon", 1.0)
    p.hllf = kw.pop("hllf", 1.0)
    p.laxf = kw.pop("laxf", 0.0)
    p.pert_amp = kw.pop("pert_amp", 4.0e-2)
    seed = kw.pop("seed", (12345, 65435, 34221, 43543))
    for k in range(4):
        p.seed[k] = seed[k]
    if kw:
        raise TypeError(f"unknown parameters: {sorted(kw)}")
    return p


def config1() -> HarmParams:
    """2D 128x128 torus, Gamma-law, HLL + MC, Courant 0.8 (the CPU-runnable case)."""
    return torus_params(128, 128, 1)


def config2() -> HarmParams:
    """3D 288x256x256 torus, Gamma-law, HLL + MC, FP64 (the headline workload)."""
    return torus_params(288, 256, 256)


def reference_default(n1: int = 256, n2: int = 256, n3: int = 1) -> HarmParams:
    """The reference's own compiled-in choices (init.c:90-131, step_ch.c:66-67)."""
    return torus_params(n1, n2, n3, a=0.6, gam=4.0 / 3.0, rin=4.0, rmax=11.8, beta=50.0,
                        cour=0.9, Rout=100.0, hllf=0.0, laxf=1.0)


def horizon_rin(a: float) -> float:
    return 0.98 * (1.0 + math.sqrt(1.0 - a * a))
