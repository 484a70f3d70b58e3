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

# Algorithmic HBM bytes per zone of each kernel class (one launch covers every
# active zone of one half-step).
#   flux: read the zone's 8 primitives, write the face's 8 fluxes.
#   emf: read 6 B-flux components, write 3 EMFs.
#   div: read the non-B fluxes of the zone's six faces (5 x 3 x 2 faces, each
#        face shared by two zones -> 5 x 3 x 1) and the three EMFs it owns,
#        write the 8 flux divergences.
#   update: read pb (8) and the divergence (8), in the corrector also pi (8);
#        write pf (8) and the flag (1 B): 193 B predictor, 257 B corrector,
#        225 B per launch on average.
ALG_BYTES = {
    "flux1": 128.0, "flux2": 128.0, "flux3": 128.0,
    "emf": 6 * 8 + 3 * 8,
    "div": 15 * 8 + 3 * 8 + 8 * 8,
    "update": ((8 + 8) * 8 + 8 * 8 + 1 + (8 + 8 + 8) * 8 + 8 * 8 + 1) / 2,
}


def cpu_baseline(threads: int) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from harm_cool_amd.params import torus_params
    from oracle_lib import Oracle, lib

    lib().orc_set_threads(threads)
    P = torus_params(288, 256, 32)
    o = Oracle(P)
    o.init_torus()
    o.step()
    nsteps = 3
    t0 = time.perf_counter()
    for _ in range(nsteps):
        o.step()
    el = time.perf_counter() - t0
    return {"value": P.n1 * P.n2 * P.n3 * nsteps / el, "unit": "zone-cycles/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle (C restatement, OpenMP) on a 288x256x32 slab of the same torus, {nsteps} steps "
                      f"after 1 warm-up, {el:.1f} s"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=str, default="288x256x256", help="grid N1xN2xN3")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the whole domain lives on one MI355X; --gpus {args.gpus} is not "
              f"supported (see DESIGN.md for how the step would shard).", file=sys.stderr)
        sys.exit(2)

    import harm_cool_amd as hc
    from harm_cool_amd.params import torus_params

    n1, n2, n3 = (int(x) for x in args.n.split("x"))
    P = torus_params(n1, n2, n3)
    g = hc.Harm(P, device=0)
    g.init()
    g.step_ch(args.warmup)
    g.reset_counters()
    g.time()  # synchronizes
    t0 = time.perf_counter()
    g.step_ch(args.steps)  # returns after the stream has drained
    el = time.perf_counter() - t0
    c = g.counters()
    nz = n1 * n2 * n3
    value = nz * args.steps / el

    # per-kernel timing with HIP events on the library stream (separate short run)
    g.set_timing(True)
    g.step_ch(3)
    tim = g.timing()
    g.set_timing(False)
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
    g.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(min(16, os.cpu_count() or 1))

    out = {
        "metric": "zone-cycles/s at 288x256x256; con2prim Newton iters/zone, MLP-guess vs plain",
        "value": value, "unit": "zone-cycles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded FM torus initial data)",
        "config": {"workload": f"Fishbone-Moncrief torus {n1}x{n2}x{n3}, gamma-law 13/9, a=0.9375, HLL+MC, "
                               "Courant 0.8, FP64, plain previous-step Newton guess",
                   "n1": n1, "n2": n2, "n3": n3},
        "newton_iters_per_zone": c["newton_iters"] / max(c["c2p_calls"], 1),
        "newton_kconv_per_zone": c["newton_kconv"] / max(c["c2p_calls"], 1),
        "roofline": roof, "kernels": kernels, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
