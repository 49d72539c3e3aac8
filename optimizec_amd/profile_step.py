"""Workload for rocprofv3 runs: a calibration copy of known size, then solver steps.

The copy (torch, float64, --copy-gib GiB read and the same written, 3 times) gives
the scale factor between rocprofv3's FETCH_SIZE / WRITE_SIZE counters and bytes; the
solver steps that follow are the same step as bench.py (k^4 exp initial condition,
seed 0, K0=4, CFL=0.1; Re=1e5 at N=4096, Re=1e7 at N=29400), in FP64 unless
--precision 1 (FP32 transforms) is given.

    rocprofv3 --kernel-trace --stats --output-format csv -d <dir> -o <name> -- \
        python optimizec_amd/profile_step.py --N 4096 --steps 50
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir> -o <name> -- \
        python optimizec_amd/profile_step.py --N 4096 --steps 5
"""
from __future__ import annotations

import argparse
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--copy-gib", type=float, default=1.0)
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--path", type=int, default=0)
    args = ap.parse_args()

    if args.copy_gib > 0:
        import torch

        n = int(args.copy_gib * (1 << 30) // 8)
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(f"calibration copy: {n * 8} bytes read and written per launch", flush=True)

    import optimizec_amd as oa

    Re = 1e7 if args.N > 4096 else 1e5
    s = oa.Solver(args.N, Re, 4.0, 0.1, precision=args.precision, path=args.path)
    s.init_ic(oa.IC_K4EXP, 0)
    s.start()
    s.step(args.steps)
    s.synchronize()
    print(f"solver: N={args.N} steps={args.steps} E,Z={s.energy_enstrophy()}", flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
