// adapt_visc.hpp -- the ADAPT_VISC viscosity controller of the reference driver
// (cuda/mem_cuda.cu:458-528): a PD controller on the filtered palinstrophy ratio that
// nudges the viscosity by at most 1% per diagnostic step toward a target of 20, with the
// Reynolds number clamped to [Re0 / 1000, 1000 Re0] around the N, K0 fit Re0.
#pragma once

#include <algorithm>
#include <cmath>

namespace adaptvisc {

struct State {
    double e_int = 0.0;
    double e_prev = 0.0;
    double target = 20.0;
    double palin_filt = 0.0;
    double palin_filt_1 = 0.0;
};

// reFromNK0 (mem_cuda.cu:467-474)
inline double re_from_n_k0(int N, double K0)
{
    return std::pow(10.0, 1.066240 * std::log10((double)N) - 0.083911 * std::log10(K0) + 1.286396);
}

// palinFilter2nd (mem_cuda.cu:476-490): two first-order stages, alpha = 1 / (tau + 1)
inline double palin_filter(State &A, double p_raw)
{
    const double alpha = 1.0 / (5.0 + 1.0);
    if (A.palin_filt == 0.0) {
        A.palin_filt_1 = p_raw;
        A.palin_filt = p_raw;
    } else {
        A.palin_filt_1 += alpha * (p_raw - A.palin_filt_1);
        A.palin_filt += alpha * (A.palin_filt_1 - A.palin_filt);
    }
    return A.palin_filt;
}

// adaptVisc (mem_cuda.cu:492-528).  Returns false inside the deadband (nothing changes);
// otherwise writes the new Re and visc = 1 / Re.
inline bool update(State &A, int N, double K0, double visc, double ratio, double *Re_out, double *visc_out)
{
    const double deadband = 0.001, max_frac = 0.01, Kp = 0.25, Ki = 0.0, Kd = 25.0;
    const double p = palin_filter(A, 10000.0 * ratio);
    const double e = (p - A.target) / A.target;
    if (std::fabs(e) < deadband) return false;
    if (Ki != 0.0) A.e_int += e;
    const double de = e - A.e_prev;
    A.e_prev = e;
    double u = Kp * e + Ki * A.e_int + Kd * de;
    u = std::max(-max_frac, std::min(max_frac, u));
    const double nu_new = visc * std::exp(u);
    const double Re0 = re_from_n_k0(N, K0);
    const double factor = 1000.0;
    const double Re_eff = std::max(Re0 / factor, std::min(factor * Re0, 1.0 / nu_new));
    *Re_out = Re_eff;
    *visc_out = 1.0 / Re_eff;
    return true;
}

}  // namespace adaptvisc
