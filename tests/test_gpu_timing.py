"""HIP-event kernel timing through the C ABI: every launch, every k-th launch, off.

Timing only brackets launches with events on the solver's stream, so a timed run must
also give the same state bit for bit as an untimed one.
"""
import numpy as np
import pytest

import optimizec_amd as oa

pytestmark = pytest.mark.gpu


def _solver():
    s = oa.Solver(64, 1e4, 4.0, 0.1)
    s.init_ic(oa.IC_K4EXP, 3)
    s.start()
    return s


def test_sampled_kernel_timing_counts_and_state():
    print("kernel timing N=64: every launch, every 10th, off", flush=True)
    s = _solver()
    s.set_kernel_timing(True)
    s.reset_kernel_timing()
    s.step(5)
    k = s.kernel_timing()
    assert k["k_row"]["launches"] == 5 and k["k_col"]["launches"] == 5
    assert k["k_row"]["total_ms"] > 0.0 and k["k_col"]["total_ms"] > 0.0

    s.set_kernel_timing(True, every=10)
    s.reset_kernel_timing()
    s.step(25)  # launches 1, 11 and 21 of each kernel
    k = s.kernel_timing()
    assert k["k_row"]["launches"] == 3 and k["k_col"]["launches"] == 3
    assert k["k_row"]["total_ms"] > 0.0

    s.set_kernel_timing(False)
    s.step(5)
    k = s.kernel_timing()
    assert k["k_row"]["launches"] == 3 and k["k_col"]["launches"] == 3
    timed = s.om2()
    s.close()

    u = _solver()
    u.step(35)
    assert np.array_equal(u.om2(), timed)
    u.close()
