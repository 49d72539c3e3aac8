"""The matrix-core variants of the mixed-precision step against the FP64 oracle, through the
C ABI.  PREC_BF16X3 runs every radix stage (radix <= 16) of the FP32 line transforms as a
DFT-matrix product with three BF16 products; PREC_FP16 as one FP16 product with each
butterfly's inputs scaled by a power of two.  State, CN/AB2, CFL/dt and E/Z stay FP64.

Bars: the nonlinear term of one step (fnm1 = FN) carries the transforms' own error
(BF16x3 <= 2e-4, FP16 <= 5e-3 relative L2; both above 1e-8, so the matrix cores really
ran).  The vorticity after 100 steps: BF16x3 within the mixed-precision bar 1e-5 of the
oracle (measured 1.4e-7 at N=512, 2.0e-8 at N=4096); FP16 within 3e-5 (measured 1.2e-5 at
N=512, 5.8e-6 at N=1024 and 2.0e-6 at N=4096: it meets the 1e-5 bar from N=1024 up).
"""
import numpy as np
import pytest

import optimizec_amd as oa

pytestmark = pytest.mark.gpu

# (precision, name, bar on FN after one step, bar on omega after 100 steps)
VARIANTS = [(oa.PREC_BF16X3, "bf16x3", 2e-4, 1e-5), (oa.PREC_FP16, "fp16", 5e-3, 3e-5)]


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.mark.parametrize("prec,name,fn_bar,om_bar", VARIANTS)
@pytest.mark.parametrize("N,path", [(256, oa.PATH_FUSED), (512, oa.PATH_FUSED), (1024, oa.PATH_FUSED),
                                    (4096, oa.PATH_FUSED), (1024, oa.PATH_MULTIPASS)])
def test_mfma_single_step_nonlinear_term(oracle_mod, prec, name, fn_bar, om_bar, N, path):
    print(f"{name} N={N} path={path}: one step", flush=True)
    s = oa.Solver(N, 1e4, 4.0, 0.1, precision=prec, path=path)
    s.init_ic(oa.IC_K4EXP, 1)
    o = oracle_mod.Oracle(N, 1e4, 4.0, 0.1)
    o.init_config(1)
    s.start()
    o.start(form=1)
    assert abs(s.scalars().dt / o.scalars().dt - 1) < 1e-14  # NEXTDT runs in FP64
    s.step(1)
    o.one_step(form=1)
    efn = relerr(s.fnm1(), o.fnm1())
    eom = relerr(s.om2(), o.om2())
    print(f"{name} N={N} path={path}: FN rel L2 {efn:.3e}, omega rel L2 {eom:.3e}")
    assert 1e-8 < efn < fn_bar
    assert eom < 1e-6
    s.close()
    o.close()


@pytest.mark.parametrize("prec,name,fn_bar,om_bar", VARIANTS)
@pytest.mark.parametrize("N", [512, 1024])
def test_mfma_100_steps_against_oracle(oracle_mod, prec, name, fn_bar, om_bar, N):
    steps, update = 100, 20
    print(f"{name} N={N}: {steps} steps", flush=True)
    s = oa.Solver(N, 1e4, 4.0, 0.1, precision=prec)
    s.init_ic(oa.IC_K4EXP, 0)
    o = oracle_mod.Oracle(N, 1e4, 4.0, 0.1)
    o.init_config(0)
    om0 = o.om2()
    s.start()
    o.start(form=1)
    rs = s.run(steps, update)
    ro = o.run(steps, update, form=1)
    om, oo = s.om2(), o.om2()
    err = relerr(om, oo)
    derr = float(np.linalg.norm(om - oo) / np.linalg.norm(oo - om0))
    serr = float(np.max(np.abs(rs[:, 5:] - ro[:, 5:]) / np.abs(ro[:, 5:])))
    print(f"{name} N={N} {steps} steps: omega rel L2 {err:.3e} (vs |w100 - w0|: {derr:.3e}), E/Z {serr:.3e}")
    assert err < om_bar
    assert serr < 1e-5
    s.close()
    o.close()


@pytest.mark.parametrize("prec,name,fn_bar,om_bar", VARIANTS)
def test_mfma_4096_config_against_oracle(oracle_4096, prec, name, fn_bar, om_bar):
    """The N=4096 configuration (Re=1e5, K0=4, CFL=0.1, seed 0), 100 steps, against the
    oracle: omega within 1e-5 (both variants); the error relative to |w_100 - w_0| is
    reported."""
    s = oa.Solver(4096, 1e5, 4.0, 0.1, precision=prec)
    s.init_ic(oa.IC_K4EXP, 0)
    s.start()
    s.run(oracle_4096["steps"], oracle_4096["update"])
    om = s.om2()
    err = relerr(om, oracle_4096["om"])
    derr = float(np.linalg.norm(om - oracle_4096["om"]) / np.linalg.norm(oracle_4096["om"] - oracle_4096["om0"]))
    print(f"{name} N=4096 100 steps: omega rel L2 {err:.3e} (vs |w100 - w0|: {derr:.3e})")
    assert err < 1e-5
    s.close()
