"""Evaluation of the learned subgrid closure against a filtered DNS.

1. An ensemble of DNS runs produces filtered training pairs (include/dns_ensemble.h).
2. The closure network is trained on them on the GPU (include/dns_closure.h).
3. A DNS at N_f = ratio * N_c and two coarse runs at N_c from the same seed (the k^4 exp
   initial condition of the coarse run is the truncation of the fine one) advance to
   the same time with a fixed time step (dt_c = ratio * dt_f, the DNS's initial CFL
   step):
   one coarse run plain (the fallback), one with the closure attached (band
   N_c/8 .. N_c/4, dissipative limiter on), one with the raw network forcing (band
   only).
4. Reported: the shell energy spectra of the coarse runs against the spectrally
   truncated DNS (relative error per shell k = 1 .. N_c/2 - 1: rms, max, max over the
   resolved shells -- those carrying at least 1e-12 of the peak shell's energy -- and max
   over the tail beyond them), the enstrophy history, and the GPU time of each run.  --dns-cache keeps the DNS reference in an
   .npz file (computed once, reused by later evaluations of the same workload).

    python -m optimizec_amd.closure_eval --coarse 1024 --ratio 8 --Re 1e6 --K0 8 \\
        --coarse-steps 5000 --ensemble config

Prints one JSON line.  The ensemble "config" is the 192 x N=256 + 64 x N=512 workload
(2000 steps, pairs every 20); "small" is 64 runs at N = 8 N_c for quick checks.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

import optimizec_amd as oa


RESOLVED_FLOOR = 1e-12


def shell_spectrum(om: np.ndarray, N: int) -> np.ndarray:
    """E(k) = sum over |k| in [k - 1/2, k + 1/2) of m |omega-hat|^2 / |k|^2 (m = 1 for kx = 0,
    else 2: the unit-energy convention of the initial condition)."""
    H = N // 2
    kx = np.arange(H)[None, :]
    kz = np.where(np.arange(N) < H, np.arange(N), np.arange(N) - N)[:, None]
    k = np.sqrt(kx ** 2 + kz ** 2)
    m = np.where(kx == 0, 1.0, 2.0)
    e = m * np.abs(om) ** 2 / np.where(k > 0, k ** 2, 1.0)
    e[0, 0] = 0.0
    return np.bincount(np.rint(k).astype(int).ravel(), e.ravel(), minlength=N)


def enstrophy(om: np.ndarray) -> float:
    """Z = sum m |omega-hat|^2 (m = 1 for kx = 0, else 2)."""
    a = np.abs(om) ** 2
    return float(np.sum(a[:, 0]) + 2.0 * np.sum(a[:, 1:]))


def truncate(om: np.ndarray, N: int, n: int) -> np.ndarray:
    """Sharp spectral filter to |kx|, |kz| < n/2, in the n-grid layout."""
    h = n // 2
    out = np.zeros((n, h), dtype=np.complex128)
    for kz in range(-h + 1, h):
        out[kz % n] = om[kz % N, :h]
    return out


def ensemble_members(kind: str, n_coarse: int, ratio: int):
    rng = np.random.default_rng(12345)
    out = []
    if kind == "config":
        for seed in range(256):
            out.append((256 if seed < 192 else 512, float(10 ** rng.uniform(3, 5)),
                        float(rng.choice([4, 8, 12, 16])), 0.1, seed))
        return out, 2000
    for i in range(64):
        out.append((n_coarse * ratio, float(10 ** rng.uniform(3.5, 5)), float(rng.choice([4, 8, 12, 16])),
                    0.1, 5000 + i))
    return out, 1200


def evaluate(coarse: int = 64, ratio: int = 8, Re: float = 1e5, K0: float = 8.0, seed: int = 1000,
             coarse_steps: int = 357, ensemble: str = "small", train_iters: int = 1500, batch: int = 64,
             band=None, samples: int = 5, log=None, dns_cache: str = None, nu_dt_max: float = 0.5) -> dict:
    def say(msg):
        if log:
            print(msg, file=log, flush=True)

    Nc, Nf = coarse, coarse * ratio
    band = band if band is not None else (Nc / 8.0, Nc / 4.0)
    members, ens_steps = ensemble_members(ensemble, Nc, ratio)
    t0 = time.perf_counter()
    e = oa.Ensemble(members, pair_every=20, filter_ratio=8, max_pairs=ens_steps // 20)
    e.init()
    e.advance(ens_steps)
    e.synchronize()
    t_ens = time.perf_counter() - t0
    say(f"ensemble: {len(members)} runs x {ens_steps} steps in {t_ens:.1f} s")
    net = oa.ClosureNet(seed=3)
    t0 = time.perf_counter()
    hist = net.train_on_ensemble(e, train_iters, batch=batch, lr=1e-3, seed=1)
    t_train = time.perf_counter() - t0
    e.close()
    say(f"training: {train_iters} Adam steps in {t_train:.1f} s, loss {hist[:20].mean():.4f} -> {hist[-50:].mean():.4f}")

    every = max(1, coarse_steps // samples)
    key = f"{Nc}_{ratio}_{Re:g}_{K0:g}_{seed}_{coarse_steps}_{samples}"
    cached = None
    if dns_cache and os.path.exists(dns_cache):
        d = np.load(dns_cache)
        if str(d["key"]) == key:
            cached = d
    if cached is not None:
        dt_f, t_end, t_dns = float(cached["dt_f"]), float(cached["t_end"]), float(cached["t_dns"])
        E_dns = cached["E_dns"]
        z_dns = [(int(a), float(b)) for a, b in cached["z_dns"]]
        say(f"DNS reference from {dns_cache}")
    else:
        # DNS, fixed step
        f = oa.Solver(Nf, Re, K0, 0.1)
        f.init_ic(oa.IC_K4EXP, seed)
        f.start()
        dt_f = f.scalars().dt
        z_dns = []
        f.synchronize()
        t0 = time.perf_counter()
        for s in range(1, coarse_steps + 1):
            f.step(ratio)
            if s % every == 0 or s == coarse_steps:
                f.synchronize()
                z_dns.append((s, enstrophy(truncate(f.om2(), Nf, Nc))))
                say(f"DNS N={Nf}: coarse step {s}/{coarse_steps}, {time.perf_counter() - t0:.1f} s")
            elif s % 250 == 0:
                f.synchronize()
                say(f"DNS N={Nf}: coarse step {s}/{coarse_steps}, {time.perf_counter() - t0:.1f} s")
        f.synchronize()
        t_dns = time.perf_counter() - t0
        t_end = f.scalars().t
        E_dns = shell_spectrum(truncate(f.om2(), Nf, Nc), Nc)
        f.close()
        if dns_cache:
            np.savez(dns_cache, key=key, dt_f=dt_f, t_end=t_end, t_dns=t_dns, E_dns=E_dns,
                     z_dns=np.array(z_dns, dtype=np.float64))

    k = np.arange(1, Nc // 2)
    # shells that carry at least RESOLVED_FLOOR of the peak shell's energy; the rest (the
    # outermost shells of a well-resolved coarse grid, 12+ decades below the peak) is
    # reported separately as the tail
    resolved = E_dns[k] >= RESOLVED_FLOOR * np.max(E_dns[k])
    runs = {}
    for name, use, limit in (("plain", False, False), ("closure", True, True), ("closure_raw", True, False)):
        c = oa.Solver(Nc, Re, K0, 0.1)
        c.init_ic(oa.IC_K4EXP, seed)
        sc = c.scalars()
        sc.dt = dt_f * ratio
        c.set_scalars(sc)
        if use:
            c.set_closure(net)
            c.set_closure_band(*band)
            c.set_closure_limiter(limit, nu_dt_max)
        z_hist = []
        c.synchronize()
        t0 = time.perf_counter()
        for s in range(1, coarse_steps + 1):
            c.step(1)
            if s % every == 0 or s == coarse_steps:
                c.synchronize()
                z_hist.append(enstrophy(c.om2()))
        c.synchronize()
        t_run = time.perf_counter() - t0
        E = shell_spectrum(c.om2(), Nc)
        r = (E[k] - E_dns[k]) / E_dns[k]
        if not np.all(np.isfinite(E)):
            r = np.full(k.size, np.inf)
        zr = [abs(a / b[1] - 1.0) for a, b in zip(z_hist, z_dns)]
        runs[name] = {
            "seconds": round(t_run, 3),
            "spectrum_rel_err_rms": float(np.sqrt(np.mean(r * r))),
            "spectrum_rel_err_max": float(np.max(np.abs(r))),
            "spectrum_rel_err_upper_half_rms": float(np.sqrt(np.mean(r[len(r) // 2:] ** 2))),
            "spectrum_rel_err_max_resolved": float(np.max(np.abs(r[resolved]))),
            "spectrum_rel_err_max_tail": float(np.max(np.abs(r[~resolved]))) if np.any(~resolved) else 0.0,
            "spectrum_energy_weighted_err": float(np.linalg.norm(E[k] - E_dns[k]) / np.linalg.norm(E_dns[k])),
            "enstrophy_rel_err_max": float(max(zr)) if np.all(np.isfinite(zr)) else float("inf"),
            "final_energy_ratio": float(np.sum(E[1:]) / np.sum(E_dns[1:])),
            "spectrum_rel_err_at": {int(q): float(r[q - 1]) for q in (1, 2, 4, 8, Nc // 4, Nc // 2 - 2, Nc // 2 - 1)},
        }
        say(f"{name}: {json.dumps(runs[name])}")
        c.close()
    net.close()
    return {
        "workload": f"coarse N={Nc} vs {ratio}x-filtered DNS N={Nf}, Re={Re:g}, K0={K0:g}, seed {seed}, "
                    f"{coarse_steps} coarse steps (dt_c = {ratio} dt_f, fixed), t_end={t_end:.4f}",
        "ensemble": f"{len(members)} runs, {ens_steps} steps, pairs every 20 ({t_ens:.1f} s)",
        "training": {"iterations": train_iters, "batch": batch, "lr": 1e-3, "seconds": round(t_train, 2),
                     "loss_first20": float(hist[:20].mean()), "loss_last50": float(hist[-50:].mean())},
        "resolved_shells": f"k = 1 .. {int(k[resolved].max())} (E >= {RESOLVED_FLOOR:g} of the peak shell)",
        "closure_band": list(band),
        "nu_dt_max": nu_dt_max,
        "dns_seconds": round(t_dns, 3),
        "dns_from_cache": cached is not None,
        "runs": runs,
        "gpu_time_ratio_dns_over_closure_run": round(t_dns / max(runs["closure"]["seconds"], 1e-9), 1),
        "enstrophy_dns": [z for _, z in z_dns],
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--coarse", type=int, default=64)
    ap.add_argument("--ratio", type=int, default=8)
    ap.add_argument("--Re", type=float, default=1e5)
    ap.add_argument("--K0", type=float, default=8.0)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--coarse-steps", type=int, default=714)
    ap.add_argument("--ensemble", choices=("small", "config"), default="small")
    ap.add_argument("--train-iters", type=int, default=1500)
    ap.add_argument("--dns-cache", default=None)
    ap.add_argument("--band", type=float, nargs=2, default=None, help="forcing ramp k_lo k_hi (default N/8 N/4)")
    ap.add_argument("--nu-dt-max", type=float, default=0.5, help="cap of the dissipative limiter")
    args = ap.parse_args()
    out = evaluate(args.coarse, args.ratio, args.Re, args.K0, args.seed, args.coarse_steps, args.ensemble,
                   args.train_iters, band=args.band, log=sys.stderr, dns_cache=args.dns_cache,
                   nu_dt_max=args.nu_dt_max)
    print(json.dumps(out), flush=True)
    ok = all(math.isfinite(out["runs"][r]["spectrum_rel_err_rms"]) for r in ("plain", "closure"))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
