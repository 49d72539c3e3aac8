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
    return 0;
}
