"""GPU tests of the batched ensemble (include/dns_ensemble.h) through the C ABI.

* Each member is bitwise equal to a single run (Solver) with the same N, Re, K0, CFL and
  seed over the same loop (NEXTDT at it == 1 and every pair_every steps), with the HIP
  graph replay and without it.
* Training pairs equal the definition recomputed on the host: the filtered fields by
  numpy's inverse FFT of the truncated spectrum, the subgrid term from the oracle's
  nonlinear term on the fine grid and on the coarse grid.
* The subgrid term vanishes for a field whose products stay below the filter cutoff.
"""
import numpy as np
import pytest

import optimizec_amd as oa

pytestmark = pytest.mark.gpu

MEMBERS = [(128, 1e3, 4.0, 0.1, 1), (256, 3e4, 8.0, 0.1, 2), (128, 1e4, 12.0, 0.1, 3)]
EVERY, RATIO = 10, 8


def _standalone(m, steps, record=()):
    N, Re, K0, cfl, seed = m
    s = oa.Solver(N, Re, K0, cfl)
    s.init_ic(oa.IC_K4EXP, seed)
    s.start()
    states = {}
    for it in range(1, steps + 1):
        if it - 1 in record:
            states[it - 1] = s.om2()
        s.step(1)
        if it == 1 or it % EVERY == 0:
            s.next_dt()
    return s, states


@pytest.mark.parametrize("use_graph,autotune", [(1, False), (0, False), (1, True), (2, False)])
def test_members_bitwise_equal_single_runs(use_graph, autotune):
    """Blocks replayed from a graph (1), issued directly (0) or the whole advance captured as
    one graph launch (2); also with the autotuner: every launch configuration it times does
    the same arithmetic per element, so the chosen one changes nothing in the results."""
    print(f"ensemble vs single runs, graph={use_graph}, autotune={autotune}", flush=True)
    e = oa.Ensemble(MEMBERS, pair_every=EVERY, filter_ratio=RATIO, max_pairs=8, use_graph=use_graph,
                    autotune=autotune)
    e.init()
    if autotune:
        for m in range(len(MEMBERS)):
            v, ms = e.tuning(m)
            print(f"member {m} (N={MEMBERS[m][0]}): configuration {v}, ms {ms}")
            # nine configurations; the radix-16 plans (7, 8) are not timed at autotune=1
            assert len(ms) == 9 and 0 <= v < 7 and all(x > 0 for x in ms[:7]) and ms[7:] == [0.0, 0.0]
            assert ms[v] == min(ms[:7])
    e.advance(40)
    e.synchronize()
    for m, mem in enumerate(MEMBERS):
        s, _ = _standalone(mem, 40)
        a, b = e.scalars(m), s.scalars()
        for f in ("it", "t", "dt", "cn", "cnm1", "cflm", "visc"):
            assert getattr(a, f) == getattr(b, f), (m, f, getattr(a, f), getattr(b, f))
        assert np.array_equal(e.om2(m), s.om2()), m
        assert np.array_equal(e.fnm1(m), s.fnm1()), m
        s.close()
    assert e.pair_info(0) == (4, 16) and e.pair_info(1) == (4, 32)
    e.close()


@pytest.mark.parametrize("configuration", [1, 2, 3, 4, 5, 6])
def test_forced_configurations_bitwise_equal_single_runs(configuration):
    """Configurations 2 and 3 double the row threads and keep one row per workgroup
    (k_row); single runs at H = 128 .. 512 use the two-row kernel (k_row_mr); 4 and 5 take
    the opaque FFT thread index, 6 four rows per row workgroup.  All do the same arithmetic
    per element, so the states agree bit for bit."""
    members = [(256, 2e4, 8.0, 0.1, 11), (512, 5e4, 4.0, 0.1, 12)]
    print(f"ensemble configuration {configuration} vs single runs, N=256 and 512", flush=True)
    e = oa.Ensemble(members, pair_every=EVERY, filter_ratio=RATIO, max_pairs=2, configuration=configuration)
    e.init()
    assert [e.tuning(m)[0] for m in range(len(members))] == [configuration] * len(members)
    e.advance(30)
    e.synchronize()
    for m, mem in enumerate(members):
        s, _ = _standalone(mem, 30)
        assert e.scalars(m).dt == s.scalars().dt, m
        assert np.array_equal(e.om2(m), s.om2()), m
        assert np.array_equal(e.fnm1(m), s.fnm1()), m
        s.close()
    e.close()


def test_configuration_out_of_range_rejected():
    with pytest.raises(oa.DnsError):
        oa.Ensemble([(256, 1e3, 4.0, 0.1, 1)], configuration=9)


def _phys(S, n):
    full = np.zeros((n, n // 2 + 1), dtype=np.complex128)
    full[:, : n // 2] = S
    return np.fft.irfft2(full, s=(n, n)) * (n * n)


def _expected_pair(oracle_mod, om, N, n):
    hc = n // 2
    rows_c = [kz % n for kz in range(-hc + 1, hc)]
    rows_f = [kz % N for kz in range(-hc + 1, hc)]
    omc = np.zeros((n, hc), dtype=np.complex128)
    omc[rows_c] = om[rows_f, :hc]
    of = oracle_mod.Oracle(N, 1e3, 4.0, 0.1)
    of.set_om2(om)
    fnf = of.nonlinear(1)
    oc = oracle_mod.Oracle(n, 1e3, 4.0, 0.1)
    oc.set_om2(omc)
    fnc = oc.nonlinear(1)
    of.close()
    oc.close()
    tau = np.zeros((n, hc), dtype=np.complex128)
    tau[rows_c] = fnf[rows_f, :hc] - fnc[rows_c]
    kx = np.arange(hc)[None, :].astype(float)
    kz = np.where(np.arange(n) < hc, np.arange(n), np.arange(n) - n)[:, None].astype(float)
    k2 = kx ** 2 + kz ** 2
    su = np.where(kx >= 1, -kz / np.where(k2 > 0, k2, 1.0), np.where(kz != 0, -1.0 / np.where(kz != 0, kz, 1.0), 0.0))
    sw = np.where(kx >= 1, kx / np.where(k2 > 0, k2, 1.0), 0.0)
    return np.stack([_phys(omc, n), _phys(1j * su * omc, n), _phys(1j * sw * omc, n), _phys(tau, n)])


def test_pairs_match_definition(oracle_mod):
    print("ensemble training pairs vs host recomputation", flush=True)
    e = oa.Ensemble(MEMBERS, pair_every=EVERY, filter_ratio=RATIO, max_pairs=8)
    e.init()
    e.advance(30)
    e.synchronize()
    for m in (0, 1):
        N = MEMBERS[m][0]
        n = N // RATIO
        got = e.pairs(m)
        assert got.shape == (3, 4, n, n)
        s, states = _standalone(MEMBERS[m], 30, record=(0, 10, 20))
        s.close()
        for p, it in enumerate((0, 10, 20)):
            want = _expected_pair(oracle_mod, states[it], N, n)
            for ch in range(4):
                err = np.linalg.norm(got[p, ch] - want[ch]) / np.linalg.norm(want[ch])
                assert err < 1e-6, (m, it, ch, err)
            # filter idempotence: the filtered field has no modes at or above n/2
            spec = np.fft.rfft2(got[p, 0].astype(np.float64))
            assert np.max(np.abs(spec[n // 2])) < 1e-5 * np.max(np.abs(spec))
    e.close()


def test_subgrid_term_vanishes_below_half_cutoff():
    """All modes at |kx|, |kz| <= 3 and a 16-point coarse grid (cutoff 8): the products
    stay below the cutoff, the filter commutes with them, tau = 0."""
    print("ensemble tau = 0 for low modes", flush=True)
    e = oa.Ensemble([(128, 1e3, 4.0, 0.1, 0)], pair_every=EVERY, filter_ratio=RATIO, max_pairs=2)
    e.init()
    N = 128
    om = np.zeros((N, N // 2), dtype=np.complex128)
    for (kx, kz), v in {(1, 2): 0.3 + 0.1j, (3, -1): -0.2 + 0.25j, (2, 3): 0.05j, (0, 2): 0.1 + 0.07j}.items():
        om[kz % N, kx] = v
        if kx == 0:
            om[(-kz) % N, 0] = np.conj(v)
    e.set_om2(0, om)
    e.advance(EVERY)
    p = e.pairs(0)[0]
    assert np.max(np.abs(p[0])) > 0.1
    assert np.max(np.abs(p[3])) < 1e-10 * np.max(np.abs(p[0]))
    e.close()


def test_bad_configurations_rejected():
    with pytest.raises(oa.DnsError) as ei:
        oa.Ensemble([(1024, 1e3, 4.0, 0.1, 0)], filter_ratio=8)  # coarse grid 128 > 64
    assert ei.value.status == 2
    e = oa.Ensemble([(128, 1e3, 4.0, 0.1, 0)], pair_every=10, max_pairs=1)
    e.init()
    with pytest.raises(oa.DnsError):
        e.advance(15)  # not a multiple of pair_every
    e.close()


@pytest.mark.parametrize("configuration", [7, 8])
def test_radix16_plan_configurations_match_single_runs(configuration):
    """Configurations 7 and 8 run every line transform with the radix-16 plan: different
    rounding, so the members agree with the single runs (radix-8 plan) to roundoff."""
    members = [(256, 2e4, 8.0, 0.1, 11), (512, 5e4, 4.0, 0.1, 12)]
    print(f"ensemble configuration {configuration} (radix-16 plan) vs single runs", flush=True)
    e = oa.Ensemble(members, pair_every=EVERY, filter_ratio=RATIO, max_pairs=2, configuration=configuration)
    e.init()
    e.advance(30)
    e.synchronize()
    for m, mem in enumerate(members):
        s, _ = _standalone(mem, 30)
        err = float(np.linalg.norm(e.om2(m) - s.om2()) / np.linalg.norm(s.om2()))
        print(f"member {m}: rel L2 vs single run {err:.2e}")
        assert 0.0 < err < 1e-12
        assert abs(e.scalars(m).dt / s.scalars().dt - 1) < 1e-12
        s.close()
    e.close()
