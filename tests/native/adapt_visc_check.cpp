// Drives the driver's ADAPT_VISC controller (optimizec_amd/driver/adapt_visc.hpp) over the
// (N, K0, visc, ratio) calls read from stdin, one per line, and prints after each call
// "changed Re visc" with 17 significant digits.  Built by tests/test_adapt_visc_cpu.py.
#include <cstdio>

#include "adapt_visc.hpp"

int main()
{
    adaptvisc::State A;
    int N;
    double K0, visc, ratio;
    while (std::scanf("%d %lf %lf %lf", &N, &K0, &visc, &ratio) == 4) {
        double Re = 0.0, v = 0.0;
        const bool changed = adaptvisc::update(A, N, K0, visc, ratio, &Re, &v);
        std::printf("%d %.17g %.17g\n", changed ? 1 : 0, Re, v);
    }
    return 0;
}
