"""ADAPT_VISC controller of the reference driver, restated for the tests.

TEST INFRASTRUCTURE ONLY: the checker for the driver's controller
(optimizec_amd/driver/adapt_visc.hpp).  Only tests/ use it.

Follows cuda/mem_cuda.cu:458-528:
  * reFromNK0 (:467-474): log10 Re0 = 1.066240 log10 N - 0.083911 log10 K0 + 1.286396
  * palinFilter2nd (:476-490): two cascaded first-order filters, alpha = 1 / (tau + 1),
    tau = 5; the first call (filter state 0) sets both stages to the input
  * adaptVisc (:492-528): p = filter(10000 x ratio), e = (p - 20) / 20; |e| < 0.001 leaves
    the viscosity alone (and the derivative memory untouched); otherwise
    u = Kp e + Kd (e - e_prev) with Kp = 0.25, Kd = 25 (Ki = 0), clamped to +-0.01,
    nu <- nu exp(u), Re = clamp(1 / nu, Re0 / 1000, 1000 Re0), visc = 1 / Re.
Pinned by tests/test_adapt_visc_cpu.py: the deadband leaves visc unchanged, a large error
moves visc by exactly exp(+-0.01), the Re clamp holds at 1000 Re0 and Re0 / 1000, and the
filter's first call and cascade are checked on hand-computed values.
"""
import math


def re_from_n_k0(N: int, K0: float) -> float:
    return 10.0 ** (1.066240 * math.log10(N) - 0.083911 * math.log10(K0) + 1.286396)


class AdaptVisc:
    def __init__(self):
        self.e_int = 0.0
        self.e_prev = 0.0
        self.target = 20.0
        self.palin_filt = 0.0
        self.palin_filt_1 = 0.0

    def filter(self, p_raw: float) -> float:
        alpha = 1.0 / (5.0 + 1.0)
        if self.palin_filt == 0.0:
            self.palin_filt_1 = p_raw
            self.palin_filt = p_raw
        else:
            self.palin_filt_1 += alpha * (p_raw - self.palin_filt_1)
            self.palin_filt += alpha * (self.palin_filt_1 - self.palin_filt)
        return self.palin_filt

    def update(self, visc: float, N: int, K0: float, ratio: float):
        """One controller call; returns (Re, visc) after it (unchanged inside the deadband)."""
        p = self.filter(10000.0 * ratio)
        e = (p - self.target) / self.target
        if abs(e) < 0.001:
            return None
        de = e - self.e_prev
        self.e_prev = e
        u = 0.25 * e + 25.0 * de
        u = max(-0.01, min(0.01, u))
        nu_new = visc * math.exp(u)
        Re0 = re_from_n_k0(N, K0)
        Re_eff = max(Re0 / 1000.0, min(1000.0 * Re0, 1.0 / nu_new))
        return Re_eff, 1.0 / Re_eff
