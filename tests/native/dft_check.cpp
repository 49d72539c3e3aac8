// Host-side check of the register DFTs of the LDS FFT engine (optimizec_amd/csrc/fft_lds.hpp):
// every radix the stage plans use, including the prime-factor composites, against a direct
// O(R^2) DFT in long double, forward and inverse.  Built by tests/test_fft_radix_cpu.py
// with hipcc for the host only; the engine's functions are compiled host + device here.
#include <hip/hip_runtime.h>
#undef __device__
#define __device__ __host__ __attribute__((device))
#include "fft_lds.hpp"

#include <cmath>
#include <complex>
#include <cstdio>

using namespace dnsfft;

template <int R, bool INV>
static double check()
{
    cd v[R];
    std::complex<long double> x[R];
    for (int i = 0; i < R; ++i) {
        v[i] = cd{std::sin(1.3 * i + 0.2), std::cos(0.7 * i * i + 0.1)};
        x[i] = {v[i].x, v[i].y};
    }
    dft<R, INV, cd>(v);
    const long double pi = 3.14159265358979323846264338327950288L;
    double err = 0.0;
    for (int k = 0; k < R; ++k) {
        std::complex<long double> s = 0;
        for (int n = 0; n < R; ++n) {
            const long double a = (INV ? 2 : -2) * pi * (long double)((n * k) % R) / R;
            s += x[n] * std::complex<long double>(cosl(a), sinl(a));
        }
        err = std::max(err, (double)std::abs(s - std::complex<long double>(v[k].x, v[k].y)));
    }
    return err;
}

template <int R>
static void report()
{
    std::printf("%d %.3e %.3e\n", R, check<R, false>(), check<R, true>());
}

// The in-place plan (fft_lines_ip) run on the host: every butterfly of every stage through
// bfly_ip with the stage strides ip_stride, DIF from natural order to ip_pos order and DIT
// from ip_pos order to natural order, against a direct DFT in long double.
template <int L, int MAXR, bool INV, bool DIT, int S>
static void ip_stages(cd *x, const cd *tw)
{
    constexpr int NSTG = num_stages(L, MAXR);
    if constexpr (S < NSTG) {
        constexpr int s = DIT ? NSTG - 1 - S : S;
        constexpr int R = stage_radix(L, s, MAXR), ST = ip_stride(L, s, MAXR), BL = R * ST;
        for (int blk = 0; blk < L / BL; ++blk)
            for (int j = 0; j < ST; ++j) {
                cd v[R];
                for (int r = 0; r < R; ++r) v[r] = x[blk * BL + j + r * ST];
                const cd w1 = tw[j * (L / BL)], w8 = tw[(8 * j * (L / BL)) % L];
                bfly_ip<R, INV, DIT, (ST > 1), cd>(v, w1, w8);
                for (int r = 0; r < R; ++r) x[blk * BL + j + r * ST] = v[r];
            }
        ip_stages<L, MAXR, INV, DIT, S + 1>(x, tw);
    }
}

template <int L, int MAXR, bool INV, bool DIT>
static double check_ip()
{
    const long double pi = 3.14159265358979323846264338327950288L;
    static cd tw[L], x[L], in[L];
    for (int t = 0; t < L; ++t) tw[t] = cd{(double)cosl(-2 * pi * t / L), (double)sinl(-2 * pi * t / L)};
    for (int i = 0; i < L; ++i) in[i] = cd{std::sin(1.3 * i + 0.2), std::cos(0.7 * i * i + 0.1)};
    for (int k = 0; k < L; ++k) {
        if (ip_idx(L, ip_pos(L, k, MAXR), MAXR) != k) return 1e30;  // not a permutation pair
        x[DIT ? ip_pos(L, k, MAXR) : k] = in[k];
    }
    ip_stages<L, MAXR, INV, DIT, 0>(x, tw);
    double err = 0.0;
    for (int k = 0; k < L; ++k) {
        std::complex<long double> s = 0;
        for (int n = 0; n < L; ++n) {
            const long double a = (INV ? 2 : -2) * pi * (long double)(((long long)n * k) % L) / L;
            s += std::complex<long double>(in[n].x, in[n].y) * std::complex<long double>(cosl(a), sinl(a));
        }
        const cd y = x[DIT ? k : ip_pos(L, k, MAXR)];
        err = std::max(err, (double)(std::abs(s - std::complex<long double>(y.x, y.y)) / std::sqrt((long double)L)));
    }
    return err;
}

template <int L, int MAXR>
static void report_ip()
{
    std::printf("ip %d %d %.3e %.3e %.3e %.3e\n", L, MAXR, check_ip<L, MAXR, false, false>(),
                check_ip<L, MAXR, true, false>(), check_ip<L, MAXR, false, true>(), check_ip<L, MAXR, true, true>());
}

int main()
{
    report<2>();
    report<3>();
    report<4>();
    report<5>();
    report<7>();
    report<8>();
    report<10>();
    report<14>();
    report<15>();
    report<16>();
    report<21>();
    report<35>();
    // the 210-point lines of the 44100-point transform run as 15 x 14
    std::printf("plan210 %d %d %d\n", num_stages(210), stage_radix(210, 0), stage_radix(210, 1));
    // powers of two: radix 8 first by default; with radices up to 16 the fewest stages,
    // largest radix first
    std::printf("plan2048 %d %d %d %d %d\n", num_stages(2048), stage_radix(2048, 0), stage_radix(2048, 1),
                stage_radix(2048, 2), stage_radix(2048, 3));
    std::printf("plan2048r16 %d %d %d %d\n", num_stages(2048, 16), stage_radix(2048, 0, 16), stage_radix(2048, 1, 16),
                stage_radix(2048, 2, 16));
    std::printf("plan1536r16 %d %d %d %d %d\n", num_stages(1536, 16), stage_radix(1536, 0, 16),
                stage_radix(1536, 1, 16), stage_radix(1536, 2, 16), stage_radix(1536, 3, 16));
    std::printf("plan128r16 %d %d %d\n", num_stages(128, 16), stage_radix(128, 0, 16), stage_radix(128, 1, 16));
    report_ip<2048, 8>();
    report_ip<2048, 16>();
    report_ip<1536, 8>();
    report_ip<1024, 8>();
    report_ip<768, 8>();
    report_ip<210, 8>();
    report_ip<96, 8>();
    report_ip<128, 16>();
    return 0;
}
