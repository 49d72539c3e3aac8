"""The oracle's restatement of the driver's output files (oracle/outputs.py) pinned on
closed forms: PGM quantization of ramps and constant fields, the percentile clip with
planted outliers, and the energy spectrum of single Fourier modes (exact bin and energy)."""
import math

import numpy as np

import outputs  # oracle/ (test infrastructure)


def test_pgm_ramp_and_constant():
    v = np.linspace(0.0, 1.0, 1001)
    px = outputs.pgm_pixels(v)
    assert px[0] == 1 and px[-1] == 255
    # pixel k is reached at the value (k - 1.5) / 254 and kept up to (k - 0.5) / 254
    for k in (2, 77, 128, 254):
        i = int(np.ceil(((k - 1.5) / 254.0) * 1000)) + 1
        assert px[i] == k or px[i] == k + 1
    assert np.all(np.diff(px.astype(int)) >= 0)
    assert np.all(outputs.pgm_pixels(np.full(50, 3.25)) == 128)
    f = np.array([0.0, np.nan, 1.0])
    assert list(outputs.pgm_pixels(f)) == [1, 128, 255]  # non-finite values map to mid range


def test_pgm_clip_drops_planted_outliers():
    rng = np.random.default_rng(0)
    body = rng.uniform(-1.0, 1.0, 100000)
    f = np.concatenate([body, [-50.0, 60.0]])
    # without the clip the outliers set the range and the body collapses to a few grey levels
    raw = outputs.pgm_pixels(f)
    assert len(np.unique(raw[:-2])) < 10
    px = outputs.pgm_pixels(f, clipped=True)
    # with it, the 0.1% / 99.9% ranks land inside the body: it spans (almost) the full scale
    assert px[-2] == 1 and px[-1] == 255
    assert px[:-2].min() <= 3 and px[:-2].max() >= 253
    frac_sat = np.mean((px[:-2] == 1) | (px[:-2] == 255))
    assert 0.0005 < frac_sat < 0.01


def test_energy_spectrum_single_modes():
    M = 48
    x = np.arange(M)
    X, Z = np.meshgrid(x, x)  # [z][x]
    nbins = max(32, 2 * M)
    for k, q in ((3, 0), (5, 2), (0, 4)):
        u = np.cos(2 * np.pi * (k * X + q * Z) / M)
        rows = outputs.energy_spectrum(u, np.zeros_like(u))
        total = sum(e for _, e, _ in rows)
        ens = [(r, e) for r, e, c in rows if e > 1e-12 * total]  # other bins hold roundoff only
        assert len(ens) == 1
        radius = math.sqrt(k * k + q * q) / (M / 2)
        b = math.floor(radius / math.sqrt(2.0) * nbins)
        assert abs(ens[0][0] - 0.5 * math.sqrt(2.0) * (2 * b + 1) / nbins) < 1e-15
        # the R2C holds M^2/2 at (k, q) [and at (0, -q) for k = 0]; doubled for 0 < kx < M/2
        want = 2 * (M * M / 2) ** 2  # k > 0: one stored entry, doubled; k = 0: (0, q) and (0, -q)
        assert abs(ens[0][1] / want - 1) < 1e-12
    # the counts of all bins add up to every mode of the half plane but (0, 0)
    rows = outputs.energy_spectrum(np.random.default_rng(1).standard_normal((M, M)), np.zeros((M, M)))
    assert sum(c for _, _, c in rows) == M * M - 1


def test_inferno_lut_and_frames():
    lut = outputs.inferno_lut()
    assert tuple(lut[0]) == (0, 0, 4) and tuple(lut[255]) == (252, 255, 164)
    # i = 128: x = 128/255 in the (0.5, 0.75] segment, t = (x - 0.5) / 0.25 = 0.00784
    assert tuple(lut[128]) == (187, 55, 83)
    # i = 64: x = 0.25098, just past the 0.25 stop, t = 0.0039: (87.4, 15.2, 108.9)
    assert tuple(lut[64]) == (87, 15, 108)
    assert outputs.movie_scale_f(4096) == 2 and outputs.movie_scale_f(512) == 1
    assert outputs.movie_scale_f(29400) == 14
    # a constant plane is grey 128 everywhere, sigma 0 -> colour index 128
    rgb, sigma = outputs.movie_frame(np.full((12, 12), 2.0), 1)
    assert sigma == 0.0 and np.all(rgb == lut[128])
    # two-valued plane: grey 1 and 255 in equal numbers, sigma 127 -> mean +- 2.5 sigma
    # covers [-190.5, 445.5]: grey 1 -> index floor((1 + 190.5) 255 / 635 + 0.5) = 77,
    # grey 255 -> floor((255 + 190.5) 255 / 635 + 0.5) = 179
    f = np.zeros((8, 8))
    f[:, 4:] = 1.0
    rgb, sigma = outputs.movie_frame(f, 2)
    assert sigma == 127.0 and rgb.shape == (4, 4, 3)
    assert np.all(rgb[:, :2] == lut[77]) and np.all(rgb[:, 2:] == lut[179])
