"""ADAPT_VISC (reference cuda/mem_cuda.cu:458-528, 872-935) on the CPU:

* the oracle's palinstrophy ratio (oracle/dns_oracle.c, or_pal_over_ens_kmax2) against
  closed forms: a single mode gives K^2 / kmax2; two modes on columns of weight 1 and 2
  give the weighted mean; the (0, 0) mode drops out; a zero field gives 0;
* the oracle's controller (oracle/adapt_visc.py) on its branches: the first filter call,
  the filter cascade on hand-computed values, the deadband, the +-1% clamp and the Re clamp;
* the driver's controller (optimizec_amd/driver/adapt_visc.hpp, compiled with g++) against
  the oracle's on a sequence of calls through every branch.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import adapt_visc as av  # oracle/ (test infrastructure)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _single(oracle_mod, N, modes):
    o = oracle_mod.Oracle(N, 1e4, 4.0, 0.1)
    om = np.zeros((N, N // 2), dtype=np.complex128)
    for (ix, iz), v in modes.items():
        om[iz, ix] = v
    o.set_om2(om)
    r = o.pal_over_ens_kmax2()
    o.close()
    return r


def test_pal_ratio_closed_forms(oracle_mod):
    N = 64
    H = N // 2
    kmax2 = (H - 1) ** 2 + H ** 2
    # one mode: K^2 / kmax2 whatever its amplitude and column weight
    assert _single(oracle_mod, N, {(3, 5): 0.7 - 0.2j}) == pytest.approx((9 + 25) / kmax2, rel=1e-15)
    assert _single(oracle_mod, N, {(4, N - 2): 1.5j}) == pytest.approx((16 + 4) / kmax2, rel=1e-15)
    assert _single(oracle_mod, N, {(H - 1, 3): 2.0}) == pytest.approx(((H - 1) ** 2 + 9) / kmax2, rel=1e-15)
    # weight 1 on kx = 0 and kx = H - 1, weight 2 elsewhere
    a2, b2 = 0.09, 0.25
    want = (1 * 4 * a2 + 2 * 2 * b2) / ((1 * a2 + 2 * b2) * kmax2)  # (0, 2) and (1, 1)
    assert _single(oracle_mod, N, {(0, 2): 0.3, (1, 1): 0.5j}) == pytest.approx(want, rel=1e-14)
    want = (1 * ((H - 1) ** 2) * a2 + 2 * 4 * b2) / ((a2 + 2 * b2) * kmax2)  # (H-1, 0) and (2, 0)
    assert _single(oracle_mod, N, {(H - 1, 0): 0.3, (2, 0): 0.5}) == pytest.approx(want, rel=1e-14)
    # the mean mode is not part of the enstrophy sum (and has K = 0)
    assert _single(oracle_mod, N, {(3, 5): 0.7, (0, 0): 5.0}) == pytest.approx(34 / kmax2, rel=1e-15)
    assert _single(oracle_mod, N, {}) == 0.0
    assert _single(oracle_mod, N, {(0, 0): 1.0}) == 0.0


def test_controller_branches():
    N, K0 = 256, 8.0
    Re0 = av.re_from_n_k0(N, K0)
    assert math.log10(Re0) == pytest.approx(1.066240 * math.log10(256) - 0.083911 * math.log10(8) + 1.286396)
    # filter: the first call takes the input, later calls cascade with alpha = 1/6
    c = av.AdaptVisc()
    assert c.filter(30.0) == 30.0
    assert c.filter(36.0) == pytest.approx(30.0 + (31.0 - 30.0) / 6.0, rel=1e-15)
    assert c.palin_filt_1 == pytest.approx(31.0, rel=1e-15)
    # deadband: p = 20 exactly (ratio 0.002) changes nothing and keeps e_prev
    c = av.AdaptVisc()
    assert c.update(1e-4, N, K0, 0.002) is None and c.e_prev == 0.0
    # a large positive error raises visc by exactly the 1% clamp, a large negative one
    # lowers it by the same factor
    c = av.AdaptVisc()
    Re, visc = c.update(1e-4, N, K0, 0.01)       # p = 100, e = 4
    assert visc == pytest.approx(1e-4 * math.exp(0.01), rel=1e-15)
    assert Re == pytest.approx(1.0 / visc, rel=1e-15)
    c = av.AdaptVisc()
    Re, visc = c.update(1e-4, N, K0, 0.0005)     # p = 5, e = -0.75
    assert visc == pytest.approx(1e-4 * math.exp(-0.01), rel=1e-15)
    # the Re clamp around Re0
    c = av.AdaptVisc()
    Re, visc = c.update(1e-12, N, K0, 0.0005)
    assert Re == pytest.approx(1000.0 * Re0, rel=1e-15) and visc == pytest.approx(1.0 / (1000.0 * Re0), rel=1e-15)
    c = av.AdaptVisc()
    Re, _ = c.update(10.0, N, K0, 0.01)
    assert Re == pytest.approx(Re0 / 1000.0, rel=1e-15)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_driver_controller_matches_oracle(tmp_path):
    exe = str(tmp_path / "adapt_visc_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", f"-I{ROOT}/optimizec_amd/driver",
                    os.path.join(ROOT, "tests", "native", "adapt_visc_check.cpp"), "-o", exe], check=True)
    rng = np.random.default_rng(7)
    calls = []
    visc = 1e-4
    ref = av.AdaptVisc()
    lines = []
    for k in range(60):
        # ratios around the target 0.002 (deadband, small and clamped corrections), a few far off
        ratio = 0.002 * (1.0 + (rng.standard_normal() * (0.002 if k % 7 == 0 else 0.2)))
        if k in (20, 41):
            ratio = 0.05
        calls.append((256, 8.0, visc, ratio))
        out = ref.update(visc, 256, 8.0, ratio)
        lines.append(out)
        if out is not None:
            visc = out[1]
    text = "".join(f"{n} {k0!r} {v!r} {r!r}\n" for n, k0, v, r in calls)
    res = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
    changed = 0
    for want, got in zip(lines, res):
        f = got.split()
        if want is None:
            assert f[0] == "0"
        else:
            changed += 1
            assert f[0] == "1"
            assert float(f[1]) == pytest.approx(want[0], rel=1e-14)
            assert float(f[2]) == pytest.approx(want[1], rel=1e-14)
    assert 10 < changed < 60  # both branches exercised
