"""Benchmark of the FP64 time step at N=4096 (6144^2 padded grid) on one MI355X.

One "step" is one pass of the mem_cuda loop body (mem_cuda.cu:1848-1857): the row
kernel (x-inverse of u, w, grad omega -> u.grad(omega) -> x-forward) and the column
kernel (z-forward -> CN/AB2 update -> z-inverse of the new omega), with NEXTDT on the
reference cadence (it % 100 == 0).  The state is resident in HBM before the timed
region; the initial condition is the seeded k^4 exp(-2 k^2/K0^2) spectrum.

Prints one JSON line (see the keys below).  ``--gpus`` above 1 is refused: this build
runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
UPDATE = 100

WORKLOAD = {"N": 4096, "Re": 1.0e5, "K0": 4.0, "CFL": 0.1, "seed": 0}


def _traffic_from_profile(kernel: str, N: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    (profiles/hbm_traffic_N{N}.json), or None when it has not been measured."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_N{N}.json")
    if not os.path.exists(path):
        return None
    with open(path, encoding="utf-8") as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else k.get("hbm_bytes_per_launch")


def cpu_baseline(N: int, seconds: float = 15.0):
    """The FP64 CPU oracle (advective form, OpenMP) on a bounded sample of the same
    workload: init + start, then as many steps as fit in about `seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # noqa: E402  (checker only: the CPU baseline leg)

    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    o = oracle.Oracle(N, WORKLOAD["Re"], WORKLOAD["K0"], WORKLOAD["CFL"])
    o.init_config(WORKLOAD["seed"])
    o.start(form=1)
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end and len(times) < 50:
        t0 = time.perf_counter()
        o.one_step(form=1)
        times.append(time.perf_counter() - t0)
    o.close()
    times.sort()
    med = times[len(times) // 2]
    return {"value": round(1e3 * med, 3), "unit": "ms/step", "cores": threads, "kind": "port",
            "sample": f"N={N} FP64 advective-form oracle, {len(times)} steps after init, median"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=WORKLOAD["N"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp64-only", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no "
              "multi-GPU path.", file=sys.stderr)
        return 2

    import numpy as np

    import optimizec_amd as oa

    N = args.N

    def make(prec):
        s = oa.Solver(N, WORKLOAD["Re"], WORKLOAD["K0"], WORKLOAD["CFL"], precision=prec)
        s.init_ic(oa.IC_K4EXP, WORKLOAD["seed"])
        s.start()
        return s

    def advance(s, n):
        for _ in range(n):
            s.step(1)
            it = s.scalars().it
            if it == 1 or it % UPDATE == 0:
                s.next_dt()

    def timed(s):
        advance(s, args.warmup)
        s.synchronize()
        s.set_kernel_timing(True)
        s.reset_kernel_timing()
        s.synchronize()
        t0 = time.perf_counter()
        advance(s, args.steps)
        s.synchronize()
        w = time.perf_counter() - t0
        k = s.kernel_timing()
        s.set_kernel_timing(False)
        return w, k

    s = make(oa.PREC_FP64)
    wall, kt = timed(s)
    E, Z = s.energy_enstrophy()
    finite = math.isfinite(E) and math.isfinite(Z)
    sc = s.scalars()
    dev_bytes = s.device_bytes()
    s.close()

    # mixed precision: FP32 transforms, FP64 state / CN-AB2 / CFL-dt / E, Z
    mixed = None
    if not args.fp64_only:
        s32 = make(oa.PREC_FP32)
        wall32, kt32 = timed(s32)
        s32.close()
        a, b = make(oa.PREC_FP64), make(oa.PREC_FP32)
        advance(a, 100)
        advance(b, 100)
        oa64, oa32 = a.om2(), b.om2()
        err = float(np.linalg.norm(oa32 - oa64) / np.linalg.norm(oa64))
        a.close()
        b.close()
        ms32 = 1e3 * wall32 / args.steps
        mixed = {"precision": "fp32 transforms + fp64 state", "ms_per_step": round(ms32, 4),
                 "speedup_vs_fp64": round((1e3 * wall / args.steps) / ms32, 3),
                 "rel_l2_vorticity_err_vs_fp64_step100": err,
                 "kernels": {n: round(d["total_ms"] / max(1, d["launches"]), 5) for n, d in kt32.items()
                             if d["launches"]}}

    ms_per_step = 1e3 * wall / args.steps
    per_kernel = {}
    for name, d in kt.items():
        if d["launches"] == 0:
            continue
        avg_ms = d["total_ms"] / d["launches"]
        per_kernel[name] = {"avg_ms": round(avg_ms, 5), "launches": d["launches"],
                            "algorithmic_GBps": round(d["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9, 1)}
    dom = max(("k_row", "k_col"), key=lambda k: kt[k]["total_ms"])
    dom_avg_ms = kt[dom]["total_ms"] / max(1, kt[dom]["launches"])
    achieved = kt[dom]["bytes_per_launch"] / (dom_avg_ms * 1e-3) / 1e9
    step_bytes = 152.0 * N * N  # 2.5 c_s + 7.5 c_i with the om2 re-read fused away
    out = {
        "metric": "step time at N=4096 & 29400, FP64 vs mixed-prec at stated L2 err; HBM GB/s",
        "value": round(ms_per_step, 4),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "steps_per_s": round(1e3 / ms_per_step, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded k^4 exp(-2 k^2/K0^2) vorticity spectrum, unit energy)",
        "config": {"workload": f"2D vorticity DNS step, N={N} ({3 * N // 2}^2 padded), FP64, Re=1e5, "
                               f"K0=4, CFL=0.1, NEXTDT every {UPDATE} steps",
                   "N": N, "padded": 3 * N // 2, "Re": WORKLOAD["Re"], "K0": WORKLOAD["K0"],
                   "CFL": WORKLOAD["CFL"], "parallelism": "single GPU"},
        "step_hbm_GBps": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
        "kernels": per_kernel,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": _traffic_from_profile(dom, N)},
        "mixed_precision": mixed,
        "device_bytes": dev_bytes,
        "final": {"it": sc.it, "t": sc.t, "dt": sc.dt, "E": E, "Z": Z, "finite": finite},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(N)
    print(json.dumps(out), flush=True)
    return 0 if finite else 1


if __name__ == "__main__":
    sys.exit(main())
