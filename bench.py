"""Throughput of the GRMHD step on one MI355X.

Workload: the Fishbone-Moncrief torus at 288x256x256 (r-theta-phi), gamma-law EOS
(Gamma=13/9), HLL flux + MC limiter, FP64, whole domain resident in HBM; one
"step" is the reference's step_ch(): predictor + corrector half-steps, each with
fluxes, flux-CT, sources, conservative update, Noble 2D inversion, floors,
boundaries and fix-ups, then dt control.

Prints one JSON line: zone-cycles/s (whole job), the roofline of the dominant
kernel (live HIP-event timing on the library's stream), Newton iterations per
zone, and the CPU oracle timed on a bounded sample of the same workload on the
host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
C5_BYTES = 64 + 128 + 128 + 8 + 64 + 64 + 3  # batched inversion, per zone (see stress())
# FP64 operations of one Noble 2D inversion counted from the restated formulas
# (add, sub, mul, div and sqrt one each; the 16-term contractions as 4 mul + 3
# add per component): setup -- U/gdet and the alpha transform 25, the scalars
# of utoprim_2d.c:240-266 (lower B, raise Q and n, B^2, Q.B, Q.n, Qt^2) 112,
# u^2 and the starting W 35, the W_last test 13, x1_of_x0 10, the primitives
# from (W, v^2) 36 -- 231; per Newton iteration func_vsq 75 (incl. 8 divisions
# and 2 square roots) and the update with its error norm 5 -- 80, counted for
# every iteration incl. the two extra ones.
C5_FLOPS_SETUP, C5_FLOPS_PER_ITER = 231, 80

# Algorithmic HBM bytes per zone of each kernel class (one launch covers every
# active zone of one half-step).
#   flux: read the zone's 8 primitives, write the face's 8 fluxes.
#   emf: read 6 B-flux components, write 3 EMFs.
#   div (the separate pass of the learned/capture/tabulated variants): read the
#        non-B fluxes of the zone's six faces (5 x 3 x 2 faces, each face shared
#        by two zones -> 5 x 3 x 1) and the three EMFs it owns, write the 8 flux
#        divergences.
#   update (the plain wave-uniform kernel, which forms the corner EMFs and the
#        divergence itself): read pb (8), the 15 non-B fluxes and the 6 B-flux
#        components of the corners the zone owns, in the corrector also pi (8);
#        write pf (8) and the flag (1 B): 297 B predictor, 361 B corrector,
#        329 B per launch on average.
ALG_BYTES = {
    "flux1": 128.0, "flux2": 128.0, "flux3": 128.0,
    "emf": 6 * 8 + 3 * 8,
    "div": 15 * 8 + 3 * 8 + 8 * 8,
    "update": ((8 + 21) * 8 + 8 * 8 + 1 + (8 + 8 + 21) * 8 + 8 * 8 + 1) / 2,
}


# One wave instruction per CU per cycle: 256 CUs at 2.4 GHz (FP64 FMA is full rate).
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 2.4e9
PROFILED_KERNEL = {"flux1": "k_flux_strip<1, false, 0>", "flux2": "k_flux_strip<2, false, 0>",
                   "flux3": "k_flux<3, true, false, 0, true>", "update": "k_update<true, 0, false, true>"}


def measured_valu(kernel: str, n1: int, n2: int, n3: int):
    """SQ_INSTS_VALU per launch of `kernel` (parity build) from the committed
    rocprofv3 counter summary of this grid, or None."""
    path = os.path.join(ROOT, "profiles", f"valu_counters_{n1}x{n2}x{n3}.txt")
    name = PROFILED_KERNEL.get(kernel)
    if not os.path.exists(path) or name is None:
        return None
    for line in open(path):
        if line.startswith("# fast build"):
            break
        if line.startswith(name):
            return float(line[len(name):].split()[0])
    return None


def measured_traffic(kernel: str, n1: int, n2: int, n3: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this grid (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel
    calibration), or None when no summary exists."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{n1}x{n2}x{n3}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


# The oracle (C restatement, OpenMP over phi planes) against the reference's
# own MPI build, both on the same 8-core Xeon (2.70 GHz) with 8 threads / 8
# ranks at 288x256x256: the reference ran 1.41e6 zone-cycles/s (steady steps,
# Makefile flags), the oracle 1.15e6-1.18e6 -- the oracle is 0.83x the
# reference there.  At 64x64x32 the reference's 8 ranks reached 3.06e5, the
# oracle 1.14e6 (the reference's halo exchanges dominate a grid that small).
ORACLE_OVER_REFERENCE_FULLSIZE = 1.165e6 / 1.41e6


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(threads: int, nsteps: int = 3) -> dict:
    """The CPU oracle on the headline grid itself (288x256x256), one warm-up
    step after init, then nsteps timed steps on `threads` host threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from harm_cool_amd.params import torus_params
    from oracle_lib import Oracle, lib

    lib().orc_set_threads(threads)
    P = torus_params(288, 256, 256)
    o = Oracle(P)
    o.init_torus()
    o.step()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        o.step()
    el = time.perf_counter() - t0
    del o
    value = P.n1 * P.n2 * P.n3 * nsteps / el
    return {"value": value, "unit": "zone-cycles/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement of the step, OpenMP over phi planes) on the same 288x256x256 torus, "
                      f"{nsteps} steps after init and 1 warm-up step, {el:.1f} s",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "per_core": value / threads,
            "oracle_over_reference_8_threads_288x256x256": round(ORACLE_OVER_REFERENCE_FULLSIZE, 3),
            "reference_equivalent_value": value / ORACLE_OVER_REFERENCE_FULLSIZE}


def stress_oracle_check(h: dict, prims: dict, gam: float, nsample: int = 4096) -> dict:
    """CPU leg: the FP64 oracle's Utoprim_2d on a sample of the stress batch,
    against the GPU's plain and learned results (max relative error)."""
    import ctypes

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import _ptr, lib

    L = lib()
    rng = np.random.default_rng(3)
    out = {}
    zs = rng.choice(h["gdet"].shape[0], nsample, replace=False)
    ref = {}
    for z in zs:
        U = np.ascontiguousarray(h["U"][:, z])
        gc = np.ascontiguousarray(h["gcov"][:, z])
        gn = np.ascontiguousarray(h["gcon"][:, z])
        p = np.ascontiguousarray(h["guess"][:, z].copy())
        ni, kc = ctypes.c_int(), ctypes.c_int()
        r = L.orc_utoprim_2d(_ptr(U), _ptr(gc), _ptr(gn), float(h["gdet"][z]), _ptr(p), gam, ctypes.byref(ni),
                             ctypes.byref(kc))
        ref[z] = (r, p)
    for name, (ret, prim) in prims.items():
        errs = []
        for z in zs:
            r, p = ref[z]
            if r == 0 and ret[z] == 0:
                errs.append(np.max(np.abs(prim[:5, z] - p[:5]) / np.maximum(np.abs(p[:5]), 1e-300)))
        e = np.array(errs)
        out[name] = {"zones_compared": int(e.size), "max_rel_err_rho_u_v": float(e.max()) if e.size else None,
                     "median_rel_err": float(np.median(e)) if e.size else None}
    return out


def capture_pairs(g, torch, nz, capture_steps):
    """Capture solver pairs over capture_steps plain steps (up to 1e7, every
    stride-th zone).  Returns (features, targets, valid, report)."""
    na = nz
    pairs = int(1e7)
    stride = max(1, int(2 * capture_steps * na / pairs))
    cap = min(pairs, 2 * capture_steps * ((na + stride - 1) // stride))
    feat = torch.zeros((8, cap), dtype=torch.float64, device="cuda:0")
    targ = torch.zeros((2, cap), dtype=torch.float64, device="cuda:0")
    valid = torch.zeros(cap, dtype=torch.uint8, device="cuda:0")
    g.set_capture(feat, targ, valid, stride)
    t0 = time.perf_counter()
    g.step_ch(capture_steps)
    t_cap = time.perf_counter() - t0
    n = g.capture_count()
    g.set_capture(None)
    fd, td, vd = feat[:, :n].contiguous(), targ[:, :n].contiguous(), valid[:n].contiguous()
    del feat, targ, valid
    rep = {"pairs": n, "capture_steps": capture_steps, "capture_stride": stride, "capture_s": round(t_cap, 2)}
    return fd, td, vd, rep


def train_seed(hc, fd, td, vd, train_steps, batch, rep):
    """Train the seed network on captured pairs. Returns (network, report)."""
    m = hc.Mlp(seed=11)
    m.fit_norm(fd, td, vd)
    m.set_loss_weights(1.0, 0.0)  # the seeded solve takes v^2 from W: train the W output only
    t0 = time.perf_counter()
    lt = m.train(fd, td, vd, batch=batch, steps=train_steps, lr0=2e-3, lr1=5e-5, trace=True)
    t_train = time.perf_counter() - t0
    rep = dict(rep, train_steps=train_steps, batch=batch, train_s=round(t_train, 2),
               loss_first=float(lt[:10].mean()), loss_last=float(lt[-50:].mean()))
    return m, rep


def stress(hc, torch, n, m_torus, train_steps, batch):
    """The inversion stress batch: plain and learned Newton starts."""
    import numpy as np

    res = {"zones": n}
    gam = 13. / 9.
    # a seed network for this distribution, trained on solver-generated pairs of a disjoint batch
    ntr = min(n, 1 << 22)
    tr = hc.stress_batch(ntr, hc.stress_spec(0, seed=(777, 888, 999, 1111)))
    ptr = tr["guess"].clone()
    rtr, _, _ = hc.utoprim2d_batch(tr["U"], tr["gcov"], tr["gcon"], tr["gdet"], ptr, gam)
    feat, targ, valid = hc.mlp_pairs(tr["U"], tr["gcov"], tr["gcon"], tr["gdet"], ptr, rtr, gam)
    del tr, ptr, rtr
    m = hc.Mlp(seed=13)
    m.fit_norm(feat, targ, valid)
    m.set_loss_weights(1.0, 0.0)
    lt = m.train(feat, targ, valid, batch=batch, steps=train_steps, lr0=2e-3, lr1=5e-5, trace=True)
    del feat, targ, valid
    res["stress_net"] = {"train_pairs": ntr, "loss_first": float(lt[:10].mean()), "loss_last": float(lt[-50:].mean())}
    sample = {}
    for kind, kname in ((0, "perturbed_1pct"), (1, "cold")):
        b = hc.stress_batch(n, hc.stress_spec(kind))
        rows = {}
        for label, net, gate in (("plain", None, 0.0), ("learned_stress_net", m, 0.0),
                                 ("learned_stress_net_gated_0.1", m, 0.1), ("learned_torus_net", m_torus, 0.0)):
            if label == "learned_torus_net" and net is None:
                continue
            prim = b["guess"].clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ret, it, kc = hc.utoprim2d_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], prim, gam, mlp=net, gate=gate)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            # algorithmic bytes per zone: U 64, gcov 128, gcon 128, gdet 8, guess in 64 and
            # result out 64, ret/iters/k_conv 3
            gbps = n * C5_BYTES / el / 1e9
            n_iter = it.double().mean().item()
            tflops = n * (C5_FLOPS_SETUP + C5_FLOPS_PER_ITER * n_iter) / el / 1e12
            rows[label] = {"n_iter_per_zone": n_iter, "k_conv_per_zone": kc.double().mean().item(),
                           "fail_rate": (ret != 0).double().mean().item(), "ms": round(el * 1e3, 2),
                           "alg_GBps": round(gbps, 1), "hbm_frac": round(gbps / HBM_PEAK_GBS, 4),
                           "fp64_TFLOPs": round(tflops, 2), "fp64_frac": round(tflops / FP64_PEAK_TFLOPS, 4)}
            if kind == 0:
                keep = slice(0, 1 << 20)
                sample[label] = (ret[keep].cpu().numpy(), prim[:, keep].cpu().numpy())
        res[kname] = rows
        if kind == 0:
            res["_host"] = {k: v[..., :1 << 20].cpu().numpy() for k, v in b.items()}
        del b
    res["_sample"] = sample
    # how good a start must be: plain Newton iterations/zone from the exact
    # primitives perturbed by a relative eps (2^22 zones of the same distribution)
    curve = []
    for eps in (0.1, 0.03, 0.01, 3e-3, 1e-3, 3e-4, 1e-4, 1e-5, 1e-6, 0.0):
        b = hc.stress_batch(1 << 22, hc.stress_spec(0, guess_pert=eps))
        prim = b["guess"].clone()
        ret, it, kc = hc.utoprim2d_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], prim, gam)
        curve.append({"eps": eps, "n_iter_per_zone": round(it.double().mean().item(), 4),
                      "fail_rate": (ret != 0).double().mean().item()})
        del b, prim
    res["iterations_vs_start_error"] = curve
    m.close()
    return res


def tabulated_torus(hc, host: dict, steps: int, warmup: int) -> dict:
    """The torus on the tabulated-EOS path (COOL_EOS build): 192x192x128,
    density scale 1e11 g/cm^3, the cooling-table generator's own table over
    log rho [6, 13] x log T [9, 12] on 200 x 200 points (eos_table())."""
    import ctypes

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from harm_cool_amd.params import torus_params

    P = torus_params(192, 192, 128)
    tab = hc.DeviceEos(host, P.gam)
    g = hc.Harm(P, device=0)
    g.set_eos(tab)
    g.init()
    g.step_ch(warmup)
    g.reset_counters()
    g.time()
    t0 = time.perf_counter()
    g.step_ch(steps)
    el = time.perf_counter() - t0
    c = g.counters()
    pr = g.prims()
    g.close()
    rho = pr[0][2:194, 2:194]
    u = pr[1][2:194, 2:194]
    inside = (rho >= host["rho"][0]) & (rho <= host["rho"][-1])
    # zones whose (rho, u) the table covers (the others take the Gamma-law branch)
    k = np.clip(np.searchsorted(host["rho"], rho) - 1, 0, host["rho"].size - 2)
    a = (rho - host["rho"][k]) / (host["rho"][k + 1] - host["rho"][k])
    elo = host["e"][k, 0] + a * (host["e"][k + 1, 0] - host["e"][k, 0])
    ehi = host["e"][k, -1] + a * (host["e"][k + 1, -1] - host["e"][k, -1])
    on = inside & (u >= elo) & (u <= ehi)
    del pr, rho, u
    nz = 192 * 192 * 128
    return {"workload": "FM torus 192x192x128, tabulated EOS from the cooling-table generator (abundance() per "
                        "point; rho_unit 1e11 g/cm^3, log rho [6,13] x log T [9,12], 200x200, HdFac at r=12), "
                        "Utoprim_5d",
            "steps": steps, "zone_cycles_per_s": nz * steps / el, "ms_per_step": el / steps * 1e3,
            "n_iter_per_zone": c["newton_iters"] / max(c["c2p_calls"], 1), "fail_5d": c["fail_2d"],
            "frac_zones_in_table_rho_range": float(inside.mean()), "frac_zones_on_table": float(on.mean())}


def cache_eos_torus(hc, tab: dict, steps: int, warmup: int) -> dict:
    """The torus on the reference's own COOL_EOS pressure path: every pressure a
    temperature solve on the cooling-table cache (the generator's 200 x 200
    table as one cache for every radius) from the zone's last T; 192x192x128,
    rho_unit 1e11 g/cm^3 and U_UNIT = P_UNIT = rho_unit c^2 as the tabulated
    section.  The first step (in the warm-up) runs the first sweep of each
    direction in j order."""
    import numpy as np

    from harm_cool_amd.params import torus_params

    ax_rho, ax_T = tab["_axes"]
    hd = hc.c3_hdfac()
    desc = hc.cache_desc(ax_rho.size, ax_rho[0], ax_rho[-1], ax_T.size, ax_T[0], ax_T[-1], 1, hd, hd, d=6,
                         ztype=hc.AXIS_CUSTOM)
    values = tab["_vals"].T.reshape(1, ax_rho.size, ax_T.size, 6)
    cache = hc.Cache.create(desc, ax_rho, ax_T, np.array([hd]), values, np.ones(values.shape[:3], np.uint8))
    rho_unit = 1e11
    u_unit = rho_unit * 2.99792458e10 * 2.99792458e10
    P = torus_params(192, 192, 128)
    g = hc.Harm(P, device=0)
    g.set_eos_cache([cache], {"rho_unit": rho_unit, "u_unit": u_unit, "p_unit": u_unit})
    g.init()
    t0 = time.perf_counter()
    g.step_ch(1)
    g.time()
    first = time.perf_counter() - t0
    g.step_ch(max(warmup - 1, 0))
    g.reset_counters()
    s0 = g.eos_solves()
    g.time()
    t0 = time.perf_counter()
    g.step_ch(steps)
    g.time()
    el = time.perf_counter() - t0
    c = g.counters()
    solves = g.eos_solves() - s0
    pr = g.prims()
    T = g.last_T()
    g.close()
    cache.close()
    nz = 192 * 192 * 128
    rho = pr[0][2:194, 2:194] * rho_unit
    inside = (rho >= ax_rho[0]) & (rho <= ax_rho[-1])
    Ta = T[2:194, 2:194]
    solved = inside & (Ta > ax_T[0]) & (Ta < ax_T[-1])
    return {"workload": "FM torus 192x192x128, the reference's COOL_EOS pressure: mnewtc<1> T solve per call on "
                        "the cooling-table cache (Akima 4x4, the generator's 200x200 table as one cache), "
                        "per-zone last T, Utoprim_5d",
            "steps": steps, "zone_cycles_per_s": nz * steps / el, "ms_per_step": el / steps * 1e3,
            "first_step_ms": first * 1e3, "n_iter_per_zone": c["newton_iters"] / max(c["c2p_calls"], 1),
            "fail_5d": c["fail_2d"], "T_solves_per_zone_cycle": solves / (nz * steps),
            "frac_zones_in_cache_rho_range": float(inside.mean()),
            "frac_zones_with_T_in_cache": float(solved.mean())}


def cache_eos_cpu(tab: dict, threads: int, n3: int = 8, nsteps: int = 2) -> dict:
    """The oracle's step with the same cache EOS on a slab of the cache-EOS
    torus (192x192 x n3 phi planes; the torus is axisymmetric up to the seeded
    perturbation, so the work per plane is the full grid's), one warm-up step
    (the first, from unset temperatures) and nsteps timed."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from harm_cool_amd.params import torus_params
    from oracle_lib import Oracle, lib, orc_ceos

    ax_rho, ax_T = tab["_axes"]
    values = np.ascontiguousarray(tab["_vals"].T.reshape(1, ax_rho.size, ax_T.size, 6))
    rho_unit = 1e11
    u_unit = rho_unit * 2.99792458e10 * 2.99792458e10
    lib().orc_set_threads(threads)
    P = torus_params(192, 192, n3)
    o = Oracle(P)
    o.set_eos_cache(orc_ceos(ax_rho, ax_T, values, P.gam, rho_unit, u_unit, u_unit))
    o.init_torus()
    o.step()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        o.step()
    el = time.perf_counter() - t0
    del o
    value = 192 * 192 * n3 * nsteps / el
    return {"value": value, "unit": "zone-cycles/s", "cores": threads, "kind": "port",
            "sample": f"oracle on a 192x192x{n3} slab of the same torus and cache, {nsteps} steps after 1 warm-up "
                      f"step, {el:.1f} s"}


def fast_mode(hc, n: tuple, steps: int, warmup: int, dev_steps: int = 10) -> dict:
    """The same step built with FMA contraction (libharmcool_fast.so), reported
    beside the parity build: its step time, per-kernel times, and its deviation
    from the parity build after dev_steps steps from the same initial state
    (field-normalized max |diff| per primitive over the active zones)."""
    import numpy as np
    from harm_cool_amd.params import torus_params

    P = torus_params(*n)
    n1, n2, n3 = n
    sl = (slice(None), slice(2, n1 + 2), slice(2, n2 + 2), slice(None))
    gp = hc.Harm(P, device=0)
    gp.init()
    gp.step_ch(dev_steps)
    pp, cp = gp.prims()[sl], gp.counters()
    gp.close()
    gf = hc.Harm(P, device=0, fast=True)
    gf.init()
    gf.step_ch(dev_steps)
    pf, cf = gf.prims()[sl], gf.counters()
    dev = {name: float(np.abs(pf[v] - pp[v]).max() / np.abs(pp[v]).max())
           for v, name in enumerate(("rho", "u", "U1", "U2", "U3", "B1", "B2", "B3"))}
    torus = pp[0] > 1e-3
    rel_rho_torus = float((np.abs(pf[0] - pp[0]) / pp[0])[torus].max())
    del pp, pf
    gf.step_ch(warmup)
    gf.reset_counters()
    gf.time()
    t0 = time.perf_counter()
    gf.step_ch(steps)
    el = time.perf_counter() - t0
    c = gf.counters()
    gf.set_timing(True)
    gf.step_ch(3)
    tim = gf.timing()
    gf.set_timing(False)
    gf.close()
    kernels = {k: {"avg_ms": round(ms / cnt, 4), "launches": int(cnt)} for k, (ms, cnt) in tim.items() if cnt}
    return {"workload": f"the same torus {n1}x{n2}x{n3} with FMA contraction in every kernel (libharmcool_fast.so)",
            "steps": steps, "zone_cycles_per_s": n1 * n2 * n3 * steps / el, "ms_per_step": el / steps * 1e3,
            "newton_iters_per_zone": c["newton_iters"] / max(c["c2p_calls"], 1),
            "deviation_after_steps": dev_steps, "field_normalized_max_diff_vs_parity": dev,
            "max_rel_diff_rho_in_torus_vs_parity": rel_rho_torus,
            "counters_equal_to_parity_after_steps": cp == cf, "kernels": kernels}


def tracer_section(g, steps: int) -> dict:
    """The reference's default build advances Lagrangian tracers every step
    (TRACERSDUMP, tracers.c:298): one per zone with rho > RHOMIN, moved with
    the interpolated u^i/u^t.  Reported apart from the headline step."""
    n_live = g.tracers_init()
    g.time()
    t0 = time.perf_counter()
    g.step_ch(steps)
    el = time.perf_counter() - t0
    g.set_timing(True)
    g.step_ch(2)
    tim = g.timing()
    g.set_timing(False)
    ms, cnt = tim["tracers"]
    t0 = time.perf_counter()
    recs = g.tracers_dump()
    el_dump = time.perf_counter() - t0
    ev = g.tracers_events()
    g.tracers_stop()
    upd = ms / max(cnt, 1)
    return {"n_live_at_seed": n_live, "ms_per_step_with_tracers": el / steps * 1e3, "steps": steps,
            "update_ms": round(upd, 4), "tracer_updates_per_s": n_live / (upd * 1e-3) if upd else None,
            "dump_ms_with_host_copy": round(el_dump * 1e3, 2), "dump_records": int(recs.size),
            "removed_during_run": int(ev.size)}


def eos_table(hc, n_rho: int = 200, n_t: int = 200) -> dict:
    """The cooling-table generator (abundance.f90, one point per GPU thread) on
    the tabulated-EOS workload's grid: log rho [6, 13] g/cm^3 x log T [9, 12] K,
    HdFac of the torus's pressure maximum; the table the 192x192x128 torus runs
    on."""
    import numpy as np

    spec = hc.eos_spec(nr=n_rho, nt=n_t)
    t0 = time.perf_counter()
    host, raw = hc.fermi_eos_table(spec, raw=True)
    el = time.perf_counter() - t0
    n = n_rho * n_t
    rho = np.repeat(host["rho"] * spec.rho_unit, n_t)
    T = np.tile(host["T"], n_rho)
    return {"workload": f"cooling-table points {n_rho}x{n_t} (rho 1e6-1e13 g/cm^3, T 1e9-1e12 K, "
                        f"HdFac {hc.c3_hdfac():.4e} s), chemical equilibrium + neutrino trapping + rates per point, "
                        "FP64",
            "points": n, "seconds": round(el, 3), "points_per_s": n / el,
            "finite": bool(np.isfinite(raw).all()), "_rho": rho, "_T": T, "_vals": raw,
            "_axes": (host["rho"] * spec.rho_unit, host["T"]), "_table": host}


def cache_lookups(hc, torch, tab: dict, n: int = 1 << 22) -> dict:
    """getvalues() on the device over the generator's table held as a cache
    (one z plane, 6 values per point): n random (rho, T) lookups, the spline
    for e with de/dT (the call of the reference's T solve) and for all six."""
    import numpy as np

    ax_rho, ax_T = tab["_axes"]
    hd = hc.c3_hdfac()
    desc = hc.cache_desc(ax_rho.size, ax_rho[0], ax_rho[-1], ax_T.size, ax_T[0], ax_T[-1], 1, hd, hd, d=6,
                         ztype=hc.AXIS_CUSTOM)
    values = tab["_vals"].T.reshape(1, ax_rho.size, ax_T.size, 6)
    cache = hc.Cache.create(desc, ax_rho, ax_T, np.array([hd]), values, np.ones(values.shape[:3], np.uint8))
    rng = np.random.default_rng(5)
    x = torch.as_tensor(10 ** rng.uniform(np.log10(ax_rho[0]), np.log10(ax_rho[-1]), n), device="cuda:0")
    y = torch.as_tensor(10 ** rng.uniform(np.log10(ax_T[0]), np.log10(ax_T[-1]), n), device="cuda:0")
    zi = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    res = {"workload": f"{n} random (rho, T) lookups in the {ax_rho.size}x{ax_T.size} cooling table (Akima "
                       "spline over 4x4 grid points, FP64)"}
    for name, mask, der in (("e_and_dedT", [1, 0, 0, 0, 0, 0], True), ("all6", None, False)):
        cache.lookup(x, y, None, zi, hc.INTERP_SPLINE, mask, der)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, _, st = cache.lookup(x, y, None, zi, hc.INTERP_SPLINE, mask, der)
        el = time.perf_counter() - t0
        res[name] = {"ms": round(el * 1e3, 3), "lookups_per_s": n / el,
                     "ok": int((st == hc.CACHE_OK).sum().item())}
    # the COOL_EOS temperature solve: e of the table at a random T, solved from a
    # start 5% away (a zone's T of the previous step)
    e_true, _, _ = cache.lookup(x, y, None, zi, hc.INTERP_SPLINE, [1, 0, 0, 0, 0, 0], False)
    e = e_true[0].contiguous()
    lo, hi = ax_T[0] * (1 + 1e-6), ax_T[-1] * (1 - 1e-6)
    T0 = torch.clamp(y * (1 + 0.05 * torch.as_tensor(rng.uniform(-1, 1, n), device="cuda:0")), lo, hi)
    T = T0.clone()
    cache.solve_T(x, e, T, zi, lo, hi)  # warm-up
    T = T0.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    berr, ret = cache.solve_T(x, e, T, zi, lo, hi)
    el = time.perf_counter() - t0
    conv = ret == 1
    res["solve_T"] = {"ms": round(el * 1e3, 3), "solves_per_s": n / el, "converged": int(conv.sum().item()),
                      "max_rel_T_err_converged": (float(((T - y).abs() / y)[conv].max().item())
                                                  if bool(conv.any()) else None)}
    if bool(conv.any()):
        # the solve's own criterion is the residual in e (tolf 1e-6); where e hardly
        # changes with T (low T at high density) a small residual leaves T loose
        e_at, _, _ = cache.lookup(x, T, None, zi, hc.INTERP_SPLINE, [1, 0, 0, 0, 0, 0], False)
        r_e = ((e_at[0] - e).abs() / e.abs())[conv]
        r_T = ((T - y).abs() / y)[conv]
        q = torch.tensor([0.5, 0.999], device=r_T.device, dtype=r_T.dtype)
        res["solve_T"]["max_rel_e_residual_converged"] = float(r_e.max().item())
        res["solve_T"]["rel_T_err_converged_q50_q999"] = [float(v) for v in torch.quantile(r_T[:1 << 20], q).tolist()]
    cache.close()
    return res


def eos_table_cpu(hc, tab: dict, sample: int = 24) -> dict:
    """The reference's own abundance_ (oracle/_ref/libeosref.so) on an evenly
    strided sample of the same grid, on one core; parity of that sample."""
    import numpy as np

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    from oracle_lib import EosRef, eosref_available
    if not eosref_available():
        return None
    R = EosRef()
    idx = np.linspace(0, tab["points"] - 1, sample).astype(int)
    t0 = time.perf_counter()
    ref = np.array([R.abundance(tab["_rho"][i], tab["_T"][i], hc.c3_hdfac()) for i in idx]).T
    el = time.perf_counter() - t0
    got = tab["_vals"][:, idx]
    rel = float((np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)).max())
    return {"kind": "reference", "cores": 1, "sample": f"{sample} strided points of the grid",
            "ms_per_point": el / sample * 1e3, "points_per_s": sample / el, "max_rel_diff_gpu_vs_reference": rel}


_T0 = time.perf_counter()


def progress(what: str) -> None:
    """One line on stderr as each section of the run ends."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {what}", file=sys.stderr, flush=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=str, default="288x256x256", help="grid N1xN2xN3")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-seed and stress sections")
    ap.add_argument("--capture-steps", type=int, default=500, help="steps whose inversions feed the training set")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--train-batch", type=int, default=65536)
    ap.add_argument("--stress-n", type=int, default=1 << 26)
    ap.add_argument("--no-tabulated", action="store_true", help="skip the tabulated-EOS torus section")
    ap.add_argument("--no-tracers", action="store_true", help="skip the tracer section")
    ap.add_argument("--no-eos-table", action="store_true", help="skip the cooling-table generator section")
    ap.add_argument("--no-fast", action="store_true", help="skip the FMA-contracted (fast mode) section")
    ap.add_argument("--lim", choices=("mc", "ppm"), default="mc",
                    help="reconstruction of the headline torus (the headline workload is MC)")
    ap.add_argument("--cache-eos-only", action="store_true",
                    help="only the cache-EOS torus section (the generator's table, then the torus), one JSON line")
    ap.add_argument("--table-points", type=str, default="200x200",
                    help="cooling-table grid of the tabulated and cache-EOS sections")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the whole domain lives on one MI355X; --gpus {args.gpus} is not "
              f"supported (see DESIGN.md for how the step would shard).", file=sys.stderr)
        sys.exit(2)

    import harm_cool_amd as hc
    from harm_cool_amd.params import torus_params

    import numpy as np
    import torch

    tnr, tnt = (int(x) for x in args.table_points.split("x"))
    if args.cache_eos_only:
        tab = eos_table(hc, tnr, tnt)
        print(json.dumps({"eos_table_s": tab["seconds"],
                          "cache_eos_torus": cache_eos_torus(hc, tab, args.steps, args.warmup)}), flush=True)
        return

    n1, n2, n3 = (int(x) for x in args.n.split("x"))
    nz = n1 * n2 * n3
    lim = {"mc": 0, "ppm": 3}[args.lim]
    P = torus_params(n1, n2, n3, lim=lim)
    g = hc.Harm(P, device=0)
    g.init()
    learned_rep, m_torus, pairs = None, None, None
    if not args.no_learned and args.capture_steps > 0:
        # the capture steps (plain steps) double as the warm-up of the timed run;
        # the network is trained after the headline and its kernel timing
        pairs = capture_pairs(g, torch, nz, args.capture_steps)
        progress("capture of the learned seed's training pairs")
    g.step_ch(args.warmup)
    p_start, f_start = g.prims(), g.pflag()
    t_start = g.time()
    g.reset_counters()
    g.time()  # synchronizes
    t0 = time.perf_counter()
    g.step_ch(args.steps)  # returns after the stream has drained
    el = time.perf_counter() - t0
    c = g.counters()
    value = nz * args.steps / el
    progress(f"headline torus: {el / args.steps * 1e3:.3f} ms/step")
    p_plain = g.prims() if pairs is not None else None

    # per-kernel timing with HIP events on the library stream (separate short run
    # from the state the timed steps started at, after two untimed steps: the
    # state's copy to the device leaves the GPU idle, and its clocks take a few
    # milliseconds of work to come back)
    g.set_prims(p_start, f_start)
    g.set_time(*t_start)
    g.step_ch(2)
    g.set_timing(True)
    g.step_ch(3)
    tim = g.timing()
    g.set_timing(False)
    progress("per-kernel timing")

    if pairs is not None:
        fd, td, vd, rep = pairs
        m_torus, learned_rep = train_seed(hc, fd, td, vd, args.train_steps, args.train_batch, rep)
        del fd, td, vd, pairs
        progress("training of the learned seed")
        sl = (slice(None), slice(2, n1 + 2), slice(2, n2 + 2), slice(None))
        rho = p_plain[sl][0]
        # mode 1: every zone from the network's seed; mode 2: only zones whose first
        # Newton step from the previous-step guess exceeds the gate
        for key, mode, gate in (("always", 1, None), ("gated_0.1", 2, 0.1), ("gated_1e-3", 2, 1e-3)):
            g.set_prims(p_start, f_start)
            g.set_time(*t_start)
            g.set_guess(mode, m_torus, gate=gate)
            g.reset_counters()
            g.time()
            t0 = time.perf_counter()
            g.step_ch(args.steps)
            el_l = time.perf_counter() - t0
            cl = g.counters()
            p_l = g.prims()
            g.set_guess(0)
            dev = [float(np.abs(p_l[sl][v] - p_plain[sl][v]).max() / np.abs(p_plain[sl][v]).max()) for v in range(8)]
            rel_rho = float((np.abs(p_l[sl][0] - rho) / rho).max())
            del p_l
            learned_rep[key] = {
                "zone_cycles_per_s": nz * args.steps / el_l, "ms_per_step": el_l / args.steps * 1e3,
                "vs_plain_throughput": el / el_l,
                "n_iter_per_zone": cl["newton_iters"] / max(cl["c2p_calls"], 1),
                "k_conv_per_zone": cl["newton_kconv"] / max(cl["c2p_calls"], 1),
                "fail_2d": cl["fail_2d"], "fail_2d_plain": c["fail_2d"],
                "max_field_normalized_diff_vs_plain": max(dev), "max_rel_diff_rho_vs_plain": rel_rho,
                "bitwise_equal_to_plain": bool(max(dev) == 0.0)}
        del p_plain, rho
        progress("learned-seed torus runs")
        g.set_prims(p_start, f_start)  # the tracer section starts where the plain run did
        g.set_time(*t_start)
    del p_start, f_start

    per = {k: (ms / cnt if cnt else 0.0) for k, (ms, cnt) in tim.items()}
    dom = max(per, key=lambda k: per[k] * max(tim[k][1], 1))
    roof = None
    if dom in ALG_BYTES:
        bytes_per_launch = ALG_BYTES[dom] * nz
        achieved = bytes_per_launch / (per[dom] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_ms": round(per[dom], 4)}
    kernels = {k: {"avg_ms": round(per[k], 4), "launches": int(tim[k][1])} for k in per if tim[k][1]}
    tracers = None if args.no_tracers else tracer_section(g, min(args.steps, 10))
    g.close()
    if roof is not None:
        roof["traffic"] = measured_traffic(dom, n1, n2, n3)
        # the stencil kernels are FP64-instruction bound: their VALU issue rate
        # against one wave instruction per CU per cycle, with the instruction
        # count per launch from the committed SQ_INSTS_VALU summary
        valu = measured_valu(dom, n1, n2, n3)
        if valu is not None:
            rate = valu / (per[dom] * 1e-3)
            roof["valu"] = {"instr_per_launch": valu, "achieved_wave_instr_per_s": rate,
                            "peak_wave_instr_per_s": VALU_PEAK_WAVE_INSTR_PER_S,
                            "frac": round(rate / VALU_PEAK_WAVE_INSTR_PER_S, 4)}

    progress("tracers")
    fast = None
    if not args.no_fast and os.path.exists(hc.LIB_FAST_PATH):
        fast = fast_mode(hc, (n1, n2, n3), args.steps, args.warmup)
        progress("fast mode")

    eos_tab = None
    if not (args.no_eos_table and args.no_tabulated):
        eos_tab = eos_table(hc, tnr, tnt)
        progress("cooling table")
    tabulated = cache_eos = None
    if not args.no_tabulated:
        tabulated = tabulated_torus(hc, eos_tab["_table"], args.steps, args.warmup)
        cache_eos = cache_eos_torus(hc, eos_tab, args.steps, args.warmup)
        progress("tabulated and cache-EOS tori")
    if not args.no_eos_table:
        eos_tab["cache_lookup"] = cache_lookups(hc, torch, eos_tab)
    else:
        eos_tab = None

    st = None
    if not args.no_learned and args.stress_n > 0:
        st = stress(hc, torch, args.stress_n, m_torus, args.train_steps, args.train_batch)
        progress("inversion stress batch")
    if m_torus is not None:
        m_torus.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(min(16, os.cpu_count() or 1))
        progress("CPU baseline")
        if st is not None:
            cpu["stress_vs_oracle"] = stress_oracle_check(st["_host"], st["_sample"], 13. / 9.)
    if st is not None:
        st.pop("_host", None)
        st.pop("_sample", None)
    if eos_tab is not None:
        if not args.no_cpu:
            eos_tab["cpu_reference"] = eos_table_cpu(hc, eos_tab)
            if cache_eos is not None:
                cache_eos["cpu_baseline"] = cache_eos_cpu(eos_tab, min(16, os.cpu_count() or 1))
        for k in ("_rho", "_T", "_vals", "_axes", "_table"):
            eos_tab.pop(k)

    out = {
        "metric": "zone-cycles/s at 288x256x256; con2prim Newton iters/zone, MLP-guess vs plain",
        "value": value, "unit": "zone-cycles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded FM torus initial data)",
        "config": {"workload": f"Fishbone-Moncrief torus {n1}x{n2}x{n3}, gamma-law 13/9, a=0.9375, "
                               f"HLL+{args.lim.upper()}, "
                               "Courant 0.8, FP64, plain previous-step Newton guess",
                   "n1": n1, "n2": n2, "n3": n3},
        "newton_iters_per_zone": c["newton_iters"] / max(c["c2p_calls"], 1),
        "newton_kconv_per_zone": c["newton_kconv"] / max(c["c2p_calls"], 1),
        "newton_stopping_rule": "reference rule for both starts: |dW/W|+|dv2/v2| <= 1e-10, then 2 extra iterations",
        "learned_seed_torus": learned_rep, "stress_batch": st, "tabulated_eos_torus": tabulated, "cache_eos_torus": cache_eos,
        "roofline": roof, "kernels": kernels, "fast_mode_fma": fast, "tracers": tracers, "eos_table": eos_tab,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
