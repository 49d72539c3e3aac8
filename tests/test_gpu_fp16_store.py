"""FP16 storage of the step's intermediates (PREC_FP16_STORE) against the FP64 oracle,
through the C ABI.  The line transforms run in FP32 (the vector stages of PREC_FP32); the
z-inverted fields V and the x-forward product G are stored as FP16 pairs, each stored line
(a field's column of V, a row of G) under its own power-of-two scale.  State, CN/AB2,
CFL/dt and E/Z stay FP64.

Bars: the nonlinear term of one step (fnm1 = FN) carries the storage rounding (FP16 keeps
11 bits: relative L2 between 1e-8, so the FP16 path really ran, and 2e-3); the vorticity
after 100 steps at the N=4096 configuration within the mixed-precision bar 1e-5, E/Z within
1e-5 at every diagnostic step.  Built for N=4096 only (the unsupported sizes are rejected
in the CPU tests).
"""
import numpy as np
import pytest

import optimizec_amd as oa

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def test_fp16_store_single_step_nonlinear_term(oracle_mod):
    N = 4096
    print(f"fp16 storage N={N}: one step", flush=True)
    s = oa.Solver(N, 1e5, 4.0, 0.1, precision=oa.PREC_FP16_STORE)
    s.init_ic(oa.IC_K4EXP, 1)
    o = oracle_mod.Oracle(N, 1e5, 4.0, 0.1)
    o.init_config(1)
    s.start()
    o.start(form=1)
    assert abs(s.scalars().dt / o.scalars().dt - 1) < 1e-14  # NEXTDT runs in FP64
    s.step(1)
    o.one_step(form=1)
    efn = relerr(s.fnm1(), o.fnm1())
    eom = relerr(s.om2(), o.om2())
    print(f"fp16 storage N={N}: FN rel L2 {efn:.3e}, omega rel L2 {eom:.3e}")
    assert 1e-8 < efn < 2e-3
    assert eom < 1e-6
    s.close()
    o.close()


def test_fp16_store_4096_config_against_oracle(oracle_4096):
    """The N=4096 configuration (Re=1e5, K0=4, CFL=0.1, seed 0), 100 steps, against the
    oracle; the error relative to |w_100 - w_0| is reported."""
    s = oa.Solver(4096, 1e5, 4.0, 0.1, precision=oa.PREC_FP16_STORE)
    s.init_ic(oa.IC_K4EXP, 0)
    s.start()
    rs = s.run(oracle_4096["steps"], oracle_4096["update"])
    om = s.om2()
    ro = oracle_4096["rows"]
    err = relerr(om, oracle_4096["om"])
    derr = float(np.linalg.norm(om - oracle_4096["om"]) / np.linalg.norm(oracle_4096["om"] - oracle_4096["om0"]))
    serr = float(np.max(np.abs(rs[:, 5:] - ro[:, 5:]) / np.abs(ro[:, 5:])))
    print(f"fp16 storage N=4096 100 steps: omega rel L2 {err:.3e} (vs |w100 - w0|: {derr:.3e}), E/Z {serr:.3e}")
    assert err < 1e-5
    assert serr < 1e-5
    s.close()


def test_fp16_store_restart_and_determinism():
    """Two runs from the same state give the same bits (the exponents are recomputed from
    the data each step; nothing carries over between solvers)."""
    out = []
    for _ in range(2):
        s = oa.Solver(4096, 1e5, 4.0, 0.1, precision=oa.PREC_FP16_STORE)
        s.init_ic(oa.IC_K4EXP, 2)
        s.start()
        s.step(5)
        out.append(s.om2())
        s.close()
    assert np.array_equal(out[0], out[1])
