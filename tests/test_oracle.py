"""Pins the FP64 CPU oracle before it is trusted as the checker of the HIP path.

Each oracle function is checked against something independent of its own code:
golden vectors (numpy pocketfft), brute-force DFT, the reference's printed known-answer
values (sim_output_*_gpu.txt), closed forms (single-shell decay, the two-mode triad
formula) and invariants (Parseval, shift theorem, energy/enstrophy conservation).
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def fftvec():
    return np.load(os.path.join(GOLD, "fft_vectors.npz"))


@pytest.fixture(scope="module")
def kats():
    with open(os.path.join(GOLD, "sim_output_kats.json"), encoding="utf-8") as f:
        return json.load(f)


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


# ---------------------------------------------------------------- 1D FFT

@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 15, 16, 21, 24, 35, 49, 64, 96])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_vs_bruteforce(oracle_mod, n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert relerr(oracle_mod.fft(x, sign), oracle_mod.dft_brute(x, sign)) < 1e-14


def test_fft_vs_golden(oracle_mod, fftvec):
    sizes = sorted(int(k[2:]) for k in fftvec.files if k.startswith("x_"))
    assert 3072 in sizes and 2205 in sizes
    for n in sizes:
        x = fftvec[f"x_{n}"]
        tol = 4e-16 * (1 + math.log2(n))
        assert relerr(oracle_mod.fft(x, -1), fftvec[f"fwd_{n}"]) < tol, n
        assert relerr(oracle_mod.fft(x, +1), fftvec[f"inv_{n}"]) < tol, n


def test_fft_parseval_and_shift(oracle_mod):
    # Parseval fixes the scale; the shift theorem fixes the sign and index of twiddles.
    for n in (768, 6144 // 4, 44100 // 20, 1470):
        rng = np.random.default_rng(n)
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        X = oracle_mod.fft(x, -1)
        assert abs(np.vdot(X, X).real / n - np.vdot(x, x).real) < 1e-12 * np.vdot(x, x).real
        s = 7
        Xs = oracle_mod.fft(np.roll(x, s), -1)
        k = np.arange(n)
        assert relerr(Xs, X * np.exp(-2j * np.pi * k * s / n)) < 1e-13
        # inverse o forward = n * identity
        assert relerr(oracle_mod.fft(X, +1), n * x) < 1e-14


# ---------------------------------------------------------------- 2D transforms

def test_2d_vs_golden(oracle_mod, fftvec):
    for M in (12, 24, 48, 96):
        assert relerr(oracle_mod.r2c_2d(fftvec[f"ur_{M}"]), fftvec[f"r2c_{M}"]) < 1e-14
        assert relerr(oracle_mod.c2r_2d(fftvec[f"uc_{M}"]), fftvec[f"c2r_{M}"]) < 1e-14


def test_2d_roundtrip(oracle_mod):
    M = 96
    x = np.random.default_rng(3).standard_normal((M, M))
    assert relerr(oracle_mod.c2r_2d(oracle_mod.r2c_2d(x)), M * M * x) < 1e-14


# ---------------------------------------------------------------- PAO-k^-3 IC known answers

def _check_pao_block(log, kat):
    lines = {ln.split("=")[0].strip(): ln for ln in log.splitlines() if "=" in ln}
    for key, text in kat["pao_text"].items():
        name = text.split("=")[0].strip()
        assert lines[name] == text, (name, lines[name], text)


def test_pao_kat_512(oracle_mod, kats):
    kat = kats["sim_output_512_1001_gpu"]
    o = oracle_mod.Oracle(512, 1e4, 10.0, 0.25)
    seed, log = o.init_pao()
    assert seed == kat["pao_seed"] == 149
    _check_pao_block(log, kat)
    # [NEXTDT INIT] and the it=1 ITERATION line (reference loop, uiuj form)
    o.start(form=0)
    s = o.scalars()
    assert abs(s.cflm / kat["nextdt_init"]["CFLM"] - 1) < 3e-7
    rows = o.run(1, 100, form=0)
    it, t, dt, cn, cflm = rows[0, :5]
    one = kat["it1"]
    assert it == 1
    assert f"{t:.10f}" == f"{one['T']:.10f}"
    assert abs(dt - one["DT"]) < 1e-8
    assert abs(cn - one["CN"]) < 1e-7
    assert abs(cflm / one["CFLM"] - 1) < 3e-7
    o.om2phys()
    assert int(o.display_sigma(2)) == one["SIG"]


@pytest.mark.slow
def test_pao_kat_4096(oracle_mod, kats):
    kat = kats["sim_output_4096_501_gpu"]
    o = oracle_mod.Oracle(4096, 1e4, 10.0, 0.25)
    seed, log = o.init_pao()
    assert seed == kat["pao_seed"]
    _check_pao_block(log, kat)
    o.start(form=0)
    assert abs(o.scalars().cflm / kat["nextdt_init"]["CFLM"] - 1) < 3e-7
    rows = o.run(1, 100, form=0)
    one = kat["it1"]
    assert f"{rows[0, 1]:.10f}" == f"{one['T']:.10f}"
    assert abs(rows[0, 3] - one["CN"]) < 1e-7  # FP32 CFLM rounding x 0.2
    assert abs(rows[0, 4] / one["CFLM"] - 1) < 3e-7
    o.close()


@pytest.mark.slow
def test_pao_kat_16384_block(oracle_mod, kats):
    lib = oracle_mod.lib()
    lib.or_pao_diagnostics.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_char_p, ctypes.c_int]
    kat = kats["sim_output_16384_21_gpu"]
    buf = ctypes.create_string_buffer(8192)
    seed = lib.or_pao_diagnostics(16384, 1e6, 5.0, buf, len(buf))
    assert seed == kat["pao_seed"] == 903
    _check_pao_block(buf.value.decode(), kat)


# ---------------------------------------------------------------- closed forms

def _set_modes(o, modes):
    om = np.zeros((o.N, o.H), dtype=np.complex128)
    for (kx, kz), v in modes.items():
        iz = kz if kz >= 0 else kz + o.N
        om[iz, kx] = v
        if kx == 0 and kz != 0:
            om[(-kz) % o.N, 0] = np.conj(v)
    o.set_om2(om)
    return om


@pytest.mark.parametrize("form", [0, 1])
def test_two_mode_triad(oracle_mod, form):
    # FN_{p+q} = -(p_z q_x - p_x q_z) w_p w_q (1/|p|^2 - 1/|q|^2), FN_{p-q} similarly:
    # fixes the sign, the 1/M^2 scale, the z_spec row map and the dealiasing.
    N = 32
    o = oracle_mod.Oracle(N, 1e3, 4.0, 0.1)
    p, q = (2, 3), (1, -5)
    wp, wq = 0.7 - 0.2j, -0.3 + 0.9j
    _set_modes(o, {p: wp, q: wq})
    fn = o.nonlinear(form)

    def k2(k):
        return k[0] ** 2 + k[1] ** 2

    cross = p[1] * q[0] - p[0] * q[1]
    expect_sum = -cross * wp * wq * (1 / k2(p) - 1 / k2(q))
    expect_dif = cross * wp * np.conj(wq) * (1 / k2(p) - 1 / k2(q))
    s = (p[0] + q[0], p[1] + q[1])
    d = (p[0] - q[0], p[1] - q[1])
    got_s = fn[s[1] % N, s[0]]
    got_d = fn[d[1] % N, d[0]]
    assert abs(got_s - expect_sum) < 1e-14 * abs(expect_sum) * 10
    assert abs(got_d - expect_dif) < 1e-14 * abs(expect_dif) * 10
    # nothing else is produced
    mask = np.ones_like(fn, dtype=bool)
    mask[s[1] % N, s[0]] = mask[d[1] % N, d[0]] = False
    assert np.max(np.abs(fn[mask])) < 1e-15


@pytest.mark.parametrize("form", [0, 1])
def test_single_shell_decay(oracle_mod, form):
    # all modes on |k| = 5: FN = 0 and omega decays by (1 - a)/(1 + a), a = nu dt K^2 / 2
    N = 32
    o = oracle_mod.Oracle(N, 100.0, 4.0, 0.1)
    om0 = _set_modes(o, {(3, 4): 0.3 + 0.1j, (4, -3): -0.2j, (5, 0): 0.15, (0, 5): 0.4 - 0.25j})
    s = o.scalars()
    s.visc, s.dt, s.cnm1, s.cn = 0.01, 0.02, 0.0, 1.0
    o.set_scalars(s)
    assert np.max(np.abs(o.nonlinear(form))) < 1e-15
    if form == 0:
        # reference path needs the velocity spectrum of om2 in UC before step2b
        o.velocity_from_om2()
        o.step2a()
        o.step2b()
        o.step3()
    else:
        o.step_adv()
    a = 0.5 * 0.01 * 0.02 * 25
    assert relerr(o.om2(), om0 * (1 - a) / (1 + a)) < 1e-15


def test_conservation_and_forms(oracle_mod):
    N = 64
    o = oracle_mod.Oracle(N, 1e4, 8.0, 0.1)
    o.init_config(11)
    om = o.om2()
    fa, fu = o.nonlinear(1), o.nonlinear(0)
    assert relerr(fa, fu) < 1e-13
    kx = np.arange(o.H)[None, :]
    kz = np.where(np.arange(N) < N // 2, np.arange(N), np.arange(N) - N)[:, None]
    k2 = (kx ** 2 + kz ** 2).astype(float)
    k2[0, 0] = 1.0
    m = np.where(kx == 0, 1.0, 2.0)
    scale = np.sum(m * np.abs(fa) * np.abs(om))
    assert abs(np.sum(m * (fa * np.conj(om)).real)) < 1e-13 * scale           # enstrophy
    assert abs(np.sum(m * (fa * np.conj(om)).real / k2)) < 1e-13 * scale      # energy


def test_calcom_roundtrip(oracle_mod):
    N = 32
    o = oracle_mod.Oracle(N, 1e3, 4.0, 0.1)
    o.init_config(5)
    om = o.om2()
    o.calcom()   # om2 from the velocity spectrum written by init_config
    assert relerr(o.om2(), om) < 1e-15


# ---------------------------------------------------------------- config IC

def test_config_ic_properties(oracle_mod):
    N, K0 = 128, 4.0
    o = oracle_mod.Oracle(N, 1e4, K0, 0.1)
    o.init_config(3)
    om = o.om2()
    E, Z = o.energy_enstrophy()
    assert abs(E - 1.0) < 1e-14
    assert om[0, 0] == 0 and np.all(om[N // 2, :] == 0)
    # Hermitian kx=0 column
    assert np.array_equal(om[1:N // 2, 0], np.conj(om[N - 1:N // 2:-1, 0]))
    # shell spectrum follows (k/K0)^4 exp(-2 (k/K0)^2) per unit k: per-mode energy
    # |w|^2/K^2 = c shape(K)/(2 pi K)
    kx = np.arange(o.H)[None, :]
    kz = np.where(np.arange(N) < N // 2, np.arange(N), np.arange(N) - N)[:, None]
    K = np.sqrt(kx ** 2 + kz ** 2)
    sel = K > 0
    e_mode = np.abs(om[sel]) ** 2 / K[sel] ** 2
    shape = (K[sel] / K0) ** 4 * np.exp(-2 * (K[sel] / K0) ** 2) / (2 * np.pi * K[sel])
    keep = shape > 1e-12 * shape.max()
    ratio = e_mode[keep] / shape[keep]
    assert np.std(ratio) < 1e-12 * np.mean(ratio)
    # phases depend on the seed; amplitudes do not; a rerun is bitwise identical
    o2 = oracle_mod.Oracle(N, 1e4, K0, 0.1)
    o2.init_config(3)
    assert np.array_equal(o2.om2(), om)
    o2.init_config(4)
    assert not np.array_equal(o2.om2(), om)
    assert np.allclose(np.abs(o2.om2()), np.abs(om), rtol=1e-14, atol=0)


def test_ic_uniform_range(oracle_mod):
    vals = [oracle_mod.ic_uniform(s, kx, kz) for s in (0, 1) for kx in range(5) for kz in range(-5, 5)]
    assert all(0.0 <= v < 1.0 for v in vals)
    assert len(set(vals)) == len(vals)


# ---------------------------------------------------------------- loop logic and diagnostics

def test_first_steps_euler_then_ab2(oracle_mod):
    # step 1 uses cnm1 = 0 (forward Euler on FN); step 2 uses cnm1 = cn = 1 (AB2 weights
    # 3/2, -1/2) even though dt is unchanged (mem_cuda.cu:1848-1857, step3_cuda.cu:172).
    N = 32
    o = oracle_mod.Oracle(N, 1e3, 4.0, 0.1)
    o.init_config(9)
    o.start(form=1)
    s = o.scalars()
    dt, nu = s.dt, s.visc
    kx = np.arange(o.H)[None, :]
    kz = np.where(np.arange(N) < N // 2, np.arange(N), np.arange(N) - N)[:, None]
    a = 0.5 * nu * dt * (kx ** 2 + kz ** 2)
    om0 = o.om2()
    fn0 = o.nonlinear(1)
    o.run(1, 100, form=1)  # step 1 (diagnostic: dt becomes dt*cn afterwards)
    dt1, cn1 = o.scalars().dt, o.scalars().cn   # the dt step 2 uses
    o.run(2, 100, form=1)  # step 2
    # recompute what the loop must have done
    om1 = ((1 - a) * om0 + dt * fn0) / (1 + a)
    chk = oracle_mod.Oracle(N, 1e3, 4.0, 0.1)
    chk.set_om2(om1)
    fn1 = chk.nonlinear(1)
    a1 = 0.5 * nu * dt1 * (kx ** 2 + kz ** 2)
    om2 = ((1 - a1) * om1 + 0.5 * dt1 * 3.0 * fn1 - 0.5 * dt1 * fn0) / (1 + a1)
    assert relerr(o.om2(), om2) < 1e-14
    s2 = o.scalars()
    assert s2.cnm1 == cn1  # step 2's STEP3 copied cn (set by the it=1 NEXTDT)


def test_display_sigma_bruteforce(oracle_mod):
    N = 32
    o = oracle_mod.Oracle(N, 1e3, 4.0, 0.1)
    o.init_config(2)
    o.om2phys()
    f = o.ur(2)
    mn, mx = f.min(), f.max()
    pix = np.clip(1.0 + (f - mn) / (mx - mn) * 254.0, 1.0, 255.0).astype(int)
    assert abs(o.display_sigma(2) - pix.std()) < 1e-9
    # om2phys is the inverse transform of the padded omega spectrum
    g = np.zeros((o.M, o.NK), dtype=np.complex128)
    om = o.om2()
    g[: N // 2, : o.H] = om[: N // 2]
    g[N:, : o.H] = om[N // 2:]
    assert relerr(f, np.fft.irfft2(g, s=(o.M, o.M)) * o.M ** 2) < 1e-14


def test_cflm_bruteforce(oracle_mod):
    N = 32
    o = oracle_mod.Oracle(N, 1e3, 4.0, 0.1)
    o.init_config(4)
    o.velocity_from_om2()
    u, w = o.ur(0), o.ur(1)
    dx = 2 * np.pi / N
    assert o.cflm() == np.max(np.abs(u) / dx + np.abs(w) / dx)
