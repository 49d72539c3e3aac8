"""Output-folder files of the reference driver, restated for the tests.

TEST INFRASTRUCTURE ONLY: the checker for mem_hip's PGM and energy_spectrum.csv writers.
Only tests/ use it.

* pgm_pixels (dns_cuda.cu:643-792): min/max of the finite values; with clipping, an
  8192-bin histogram over [min, max], bin b = floor((v - min)(bins - 1)/(max - min)), the
  lo/hi ranks floor(pct (n - 1)), the first bins whose running count exceeds them, and
  [lo, hi) = [min + w lo_bin, min + w (hi_bin + 1)) with w = (max - min)/bins; pixels
  floor(1 + 254 (v - lo)/(hi - lo) + 0.5) clamped to [1, 255], a constant field (range
  <= 1e-12) 128.  The FP64 build evaluates this in double (the FP32 reference in float).
* energy_spectrum (mem_cuda.cu:575-649): the unnormalized R2C of the physical u and w on
  the M x M grid, nbins = max(32, 2M), r = |k| / (M/2), bin = floor(r / sqrt(2) nbins),
  energy |U|^2 + |W|^2 doubled and counted twice for 0 < kx < M/2, (0, 0) skipped, rows
  only where count > 0 and energy > 0, at the bin's mid radius.
Pinned by tests/test_outputs_cpu.py (ramps, constant fields, clipping of planted outliers,
single Fourier modes).
"""
import math

import numpy as np


def pgm_pixels(f, clipped=False, lo_pct=0.001, hi_pct=0.999):
    f = np.asarray(f, dtype=np.float64)
    fin = np.isfinite(f)
    n = int(fin.sum())
    mn, mx = (float(f[fin].min()), float(f[fin].max())) if n else (0.0, 0.0)
    lo, hi = mn, mx
    if clipped and n > 0 and mx - mn > 1e-12 and lo_pct < hi_pct:
        bins = 8192
        inv = (bins - 1) / (mx - mn)
        b = np.clip(((f[fin] - mn) * inv).astype(np.int64), 0, bins - 1)
        hist = np.bincount(b, minlength=bins)
        acc = np.cumsum(hist)
        lo_rank = math.floor(lo_pct * (n - 1))
        hi_rank = math.floor(hi_pct * (n - 1))
        lo_bin = int(np.argmax(acc > lo_rank))
        hi_bin = int(np.argmax(acc > hi_rank))
        w = (mx - mn) / bins
        lo, hi = mn + w * lo_bin, mn + w * (hi_bin + 1)
        if not hi > lo:
            lo, hi = mn, mx
    rng = hi - lo
    if rng <= 1e-12:
        return np.full(f.shape, 128, dtype=np.uint8)
    norm = np.where(fin, (f - lo) / rng, 0.5)
    px = np.floor(1.0 + norm * 254.0 + 0.5)
    return np.clip(px, 1, 255).astype(np.uint8)


def energy_spectrum(u, w):
    """Rows (mid radius, energy, count) of energy_spectrum.csv for physical u, w [z][x]."""
    M = u.shape[0]
    U = np.fft.rfft2(u)
    W = np.fft.rfft2(w)
    nk = M // 2 + 1
    nbins = max(32, 2 * M)
    rmax = math.sqrt(2.0)
    knyq = 0.5 * M
    esum = np.zeros(nbins)
    cnt = np.zeros(nbins, dtype=np.int64)
    z = np.arange(M)
    kz = np.where(z <= M // 2, z, z - M)[:, None]
    kx = np.arange(nk)[None, :]
    r = np.sqrt((kx * kx + kz * kz).astype(np.float64)) / knyq
    b = np.clip(np.floor(r / rmax * nbins).astype(np.int64), 0, nbins - 1)
    p = np.abs(U) ** 2 + np.abs(W) ** 2
    inner = (kx > 0) & (kx < M // 2)
    p = np.where(inner, 2.0 * p, p)
    c = np.where(inner, 2, 1) * np.ones_like(p, dtype=np.int64)
    keep = ~((kx == 0) & (kz == 0))
    np.add.at(esum, b[keep], p[keep])
    np.add.at(cnt, b[keep], c[keep])
    rows = []
    for i in range(nbins):
        if cnt[i] > 0 and esum[i] > 0.0:
            r0, r1 = rmax * i / nbins, rmax * (i + 1) / nbins
            rows.append((0.5 * (r0 + r1), float(esum[i]), int(cnt[i])))
    return rows
