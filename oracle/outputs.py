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


# ---- MOV frames (color_cuda.cu:19-597, mem_cuda.cu:245-332, 1834-1910) ----------------
def inferno_lut():
    """256 x 3 bytes: linear ramp through the stops (0, (0,0,4)), (.25, (87,15,109)),
    (.5, (187,55,84)), (.75, (249,142,8)), (1, (252,255,164)), evaluated in float32 and
    truncated, as the reference builds it (color_cuda.cu:40-70)."""
    stops = [(0.00, (0, 0, 4)), (0.25, (87, 15, 109)), (0.50, (187, 55, 84)), (0.75, (249, 142, 8)),
             (1.00, (252, 255, 164))]
    f32 = np.float32
    lut = np.zeros((256, 3), dtype=np.uint8)
    for i in range(256):
        x = f32(i) / f32(255.0)
        s = 0
        while s + 1 < 5 and x > f32(stops[s + 1][0]):
            s += 1
        (pa, ca), (pb, cb) = stops[s], stops[min(s + 1, 4)]
        span = max(f32(1e-12), f32(pb) - f32(pa))
        t = min(f32(1.0), max(f32(0.0), (x - f32(pa)) / span))
        for k in range(3):
            lut[i, k] = int((f32(1.0) - t) * f32(ca[k]) + t * f32(cb[k]))
    return lut


def movie_scale_f(N):
    """dnsCudaMovieScaleF (color_cuda.cu:486-493)"""
    return {9216: 8, 16384: 12, 29400: 14}.get(N, max(1, (3 * N // 2 + 3839) // 3840))


def movie_frame(f, F):
    """RGB frame of an M x M plane [z][x] sampled every F points (renderInfernoKernel,
    color_cuda.cu:185-233), in double; returns (rgb [h][w][3], grey-level sigma)."""
    f = np.asarray(f, dtype=np.float64)
    M = f.shape[0]
    fin = np.isfinite(f)
    mn, mx = (float(f[fin].min()), float(f[fin].max())) if fin.any() else (0.0, 0.0)
    rng = mx - mn
    const = not rng > 1e-12
    inv = 1.0 if const else 1.0 / rng
    grey = np.where(fin & (not const), np.clip(1.0 + (f - mn) * inv * 254.0, 1.0, 255.0), 128.0)
    grey = np.floor(grey).astype(np.int64)
    mean = grey.mean()
    sigma = math.sqrt(max(0.0, (grey.astype(np.float64) ** 2).mean() - mean * mean))
    idx = np.minimum(np.arange(0, (M + F - 1) // F) * F, M - 1)
    pix = grey[np.ix_(idx, idx)]
    lo, hi = mean - 2.5 * sigma, mean + 2.5 * sigma
    if sigma >= 1.0 and hi > lo:
        pix = np.clip(np.floor((pix - lo) * 255.0 / (hi - lo) + 0.5), 0, 255).astype(np.int64)
    return inferno_lut()[pix], sigma


def read_png_rgb(data):
    """(width, height, rgb) of an 8-bit RGB PNG with filter 0 rows; checks every CRC."""
    import struct
    import zlib

    assert data[:8] == bytes([137, 80, 78, 71, 13, 10, 26, 10])
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert zlib.crc32(typ + body) & 0xFFFFFFFF == crc, typ
        if typ == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert (depth, ctype) == (8, 2)
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), dtype=np.uint8).reshape(h, 3 * w + 1)
    assert np.all(raw[:, 0] == 0)
    return w, h, raw[:, 1:].reshape(h, w, 3)
