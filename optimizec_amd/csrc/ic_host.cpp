// ic_host.cpp -- initial conditions, generated on the host in FP64 (plus the FP32
// arithmetic the reference generator uses, where its results depend on it).
//
// dns_ic_pao_host: the PAO-k^-3 spectrum of pao_km3_cuda.cu:59-498 (Fortran LCG
//   phases, Hermitian kx=0 column, k^-3 shell reshaping with hash jitter/phases,
//   printed diagnostics), followed by CALCOM (calcom_cuda.cu:84-132) to om2.  The
//   values are float-representable exactly as the reference produces them, so an
//   FP64 run starts from the reference's own initial field.
// dns_ic_k4exp_host: seeded random phases with E(k) ~ (k/K0)^4 exp(-2 (k/K0)^2)
//   (no reference code): per-mode energy |u_k|^2 = c S(K) / (2 pi K), phases from a
//   counter-based generator keyed by (seed, kx, kz), Hermitian kx=0 column, the
//   kz=-N/2 row and (0,0) zero, c fixed so that sum_k m |u_k|^2 = 1 (m = 1 for
//   kx=0, 2 otherwise), i.e. the reference energy convention Q2 = 1.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/dns_hip.h"

namespace {

struct f2 {
    float re, im;
};

// Fortran-style LCG of pao_km3_cuda.cu:38-46
inline float lcg_next(int &seed)
{
    seed = (seed * 420029 + 2017) % 5011;
    return float(seed) / float(5011);
}

// sin-hash of pao_km3_cuda.cu:48-57
inline double hash01(int x, int z, int seed, int salt)
{
    const double a = (double(x) + 1.0) * 12.9898 + (double(z) + 1.0) * 78.233 +
                     (double(seed) + 1.0) * 0.010001 + (double(salt) + 1.0) * 37.719;
    const double s = std::sin(a) * 43758.5453123;
    return s - std::floor(s);
}

struct Appender {
    char *buf;
    int cap, pos = 0;
    void line(const char *fmt, double v)
    {
        if (!buf || pos >= cap) return;
        const int w = std::snprintf(buf + pos, size_t(cap - pos), fmt, v);
        if (w > 0) pos += w;
    }
    void iline(const char *fmt, int v)
    {
        if (!buf || pos >= cap) return;
        const int w = std::snprintf(buf + pos, size_t(cap - pos), fmt, v);
        if (w > 0) pos += w;
    }
};

inline uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline double uniform01(uint64_t seed, int kx, int kz)
{
    const uint64_t key = (uint64_t(uint32_t(kx)) << 32) | uint64_t(uint32_t(kz));
    return double(mix64(seed ^ mix64(key)) >> 11) * (1.0 / 9007199254740992.0);
}

// static partition of [0, n) over up to 16 host threads
template <class F>
void parallel_for(int n, F f)
{
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt == 0 ? 1 : (nt > 16 ? 16 : nt);
    if (n < 64) nt = 1;
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t) {
        const int lo = int((long long)n * t / nt), hi = int((long long)n * (t + 1) / nt);
        pool.emplace_back([=]() {
            for (int i = lo; i < hi; ++i) f(i);
        });
    }
    for (std::thread &th : pool) th.join();
}

inline double k4exp_mode_energy(double K, double K0)
{
    const double q = K / K0;
    const double q2 = q * q;
    return (q2 * q2) * std::exp(-2.0 * q2) / (2.0 * 3.14159265358979323846 * K);
}

}  // namespace

extern "C" int dns_ic_pao_host(int N, double Re, double K0d, double *om2, char *log, int logcap, double *visc_out)
{
    if (N < 4 || (N % 4) != 0 || !(Re > 0.0)) return -1;
    const int ND2 = N / 2;
    const float PI = 3.14159265358979f;
    const float DXZ = 2.0f * PI / float(N);
    const float K0 = float(K0d);
    const float NORM = PI * K0 * K0;
    const float FLOOR2 = 1.0e-40f;

    // wavenumbers (pao_km3_cuda.cu:75-92); E1 = E3 = 1
    std::vector<float> alfa(ND2), gam(N);
    for (int x = 0; x < ND2; ++x) alfa[x] = float(x);
    gam[0] = 0.0f;
    for (int z = 1; z <= ND2; ++z) {
        gam[z] = float(z);
        gam[N - z] = -gam[z];
    }

    // velocity spectrum (u1, u3) on x < N/2, z < N
    std::vector<f2> U((size_t)ND2 * N * 2, f2{0.0f, 0.0f});
    auto at = [&](int x, int z, int c) -> f2 & { return U[(size_t)c + 2 * ((size_t)x + (size_t)ND2 * z)]; };

    const int seed0 = 1;
    int seed = seed0;
    float ranvec[97];
    for (int i = 0; i < 97; ++i) lcg_next(seed);
    for (int i = 0; i < 97; ++i) ranvec[i] = lcg_next(seed);

    // isotropic random start spectrum (Fortran DO 500/510)
    for (int z = 0; z < N; ++z) {
        for (int x = 0; x < ND2; ++x) {
            const float r = lcg_next(seed);
            int slot = int(r * 97.0f);
            slot = slot < 0 ? 0 : (slot > 96 ? 96 : slot);
            const float pick = ranvec[slot];
            ranvec[slot] = r;
            const float th = 2.0f * PI * pick;
            const float cth = std::cos(th), sth = std::sin(th);
            const float ax = alfa[x], gz = gam[z];
            const float K = std::sqrt(ax * ax + gz * gz);
            const float q = K / K0;
            if (ax == 0.0f) {
                float e2 = std::exp(-q * q) / NORM;
                if (e2 < FLOOR2) e2 = FLOOR2;
                const float amp = std::sqrt(e2);
                at(x, z, 0) = f2{amp * cth, amp * sth};
                at(x, z, 1) = f2{0.0f, 0.0f};
            } else {
                const float den = 1.0f + (gz * gz) / (ax * ax);
                float e2 = std::exp(-q * q) / (den * NORM);
                if (e2 < FLOOR2) e2 = FLOOR2;
                const float amp = std::sqrt(e2);
                const f2 w{amp * cth, amp * sth};
                const float ratio = -(gz / ax);
                at(x, z, 1) = w;
                at(x, z, 0) = f2{ratio * w.re, ratio * w.im};
            }
        }
    }
    at(0, 0, 0) = at(0, 0, 1) = f2{0.0f, 0.0f};
    for (int z = 1; z < ND2; ++z)
        for (int c = 0; c < 2; ++c) at(0, N - z, c) = f2{at(0, z, c).re, -at(0, z, c).im};
    for (int x = 0; x < ND2; ++x) at(x, ND2, 0) = at(x, ND2, 1) = f2{0.0f, 0.0f};

    // k^-3 reshaping of integer shells with hash jitter and hash phases
    {
        const double knyq = double(ND2);
        std::vector<double> shell_e(ND2 + 1, 0.0), shell_w(ND2 + 1, 0.0), target(ND2 + 1, 0.0);
        for (int x = 0; x < ND2; ++x) {
            const double m = (x == 0) ? 1.0 : 2.0;
            const double ax2 = double(alfa[x]) * double(alfa[x]);
            for (int z = 0; z < N; ++z) {
                const double gz = double(gam[z]);
                const double k = std::sqrt(ax2 + gz * gz);
                const int shell = int(k + 0.5);
                if (k > 0.0 && k <= knyq && shell <= ND2) {
                    const f2 a = at(x, z, 0), b = at(x, z, 1);
                    const double ea = double(a.re) * double(a.re) + double(a.im) * double(a.im);
                    const double eb = double(b.re) * double(b.re) + double(b.im) * double(b.im);
                    const int zh = (x == 0 && z > ND2) ? N - z : z;
                    shell_e[shell] += m * (ea + eb);
                    shell_w[shell] += m * (0.75 + 0.50 * hash01(x, zh, seed0, 17));
                }
            }
        }
        const double k0 = double(K0) > 1.0 ? double(K0) : 1.0;
        double e_orig = 0.0, e_target = 0.0;
        for (int s = 1; s <= ND2; ++s) {
            const double k = double(s);
            if (k < k0) {
                const double low = k / k0;
                target[s] = (low * low * low * low) / (k0 * k0 * k0);
            } else {
                target[s] = 1.0 / (k * k * k);
            }
        }
        for (int s = 1; s <= ND2; ++s) {
            e_orig += shell_e[s];
            e_target += target[s];
        }
        if (e_orig > 0.0 && e_target > 0.0) {
            const double norm = e_orig / e_target;
            for (int x = 0; x < ND2; ++x) {
                const double ax = double(alfa[x]);
                const double ax2 = ax * ax;
                for (int z = 0; z < N; ++z) {
                    const double gz = double(gam[z]);
                    const double gz2 = gz * gz;
                    const double k = std::sqrt(ax2 + gz2);
                    const int shell = int(k + 0.5);
                    if (k > 0.0 && k <= knyq && shell <= ND2 && shell_w[shell] > 0.0) {
                        int zh = z;
                        double sgn = 1.0;
                        if (x == 0 && z > ND2) {
                            zh = N - z;
                            sgn = -1.0;
                        }
                        const double jit = 0.75 + 0.50 * hash01(x, zh, seed0, 17);
                        const double e_mode = (norm * target[shell]) * jit / shell_w[shell];
                        const double th = sgn * 2.0 * 3.14159265358979323846 * hash01(x, zh, seed0, 53);
                        const float pr = float(std::cos(th)), pi = float(std::sin(th));
                        if (ax == 0.0) {
                            const float amp = float(std::sqrt(e_mode));
                            at(x, z, 0) = f2{amp * pr, amp * pi};
                            at(x, z, 1) = f2{0.0f, 0.0f};
                        } else {
                            const float amp = float(std::sqrt(e_mode / (1.0 + gz2 / ax2)));
                            const f2 w{amp * pr, amp * pi};
                            const float ratio = float(-gz / ax);
                            at(x, z, 0) = f2{ratio * w.re, ratio * w.im};
                            at(x, z, 1) = w;
                        }
                    } else {
                        at(x, z, 0) = at(x, z, 1) = f2{0.0f, 0.0f};
                    }
                }
            }
        }
    }

    // spectral averages and the printed block (pao_km3_cuda.cu:345-424); the
    // reference squares through std::norm(complex<float>), a float sum
    double A1 = 0, A2 = 0, A3 = 0, A4 = 0, A5 = 0, A6 = 0, A7 = 0, E110 = 0;
    for (int x = 0; x < ND2; ++x) {
        const double m = (x == 0) ? 1.0 : 2.0;
        const double ax2 = double(alfa[x]) * double(alfa[x]);
        for (int z = 0; z < N; ++z) {
            const f2 a = at(x, z, 0), b = at(x, z, 1);
            const float na_re = a.re * a.re, na_im = a.im * a.im;
            const float nb_re = b.re * b.re, nb_im = b.im * b.im;
            const double uu = double(float(na_re + na_im));
            const double ww = double(float(nb_re + nb_im));
            const double gz2 = double(gam[z]) * double(gam[z]);
            const double K2 = ax2 + gz2;
            A1 += m * uu;
            A2 += m * ww;
            A3 += m * uu * ax2;
            A4 += m * uu * gz2;
            A5 += m * ww * ax2;
            A6 += m * ww * gz2;
            A7 += m * (uu + ww) * K2 * K2;
            if (x == 0) E110 += uu;
        }
    }
    const double Q2 = A1 + A2, W2 = A3 + A4 + A5 + A6;
    const double visc = 1.0 / Re;
    if (visc_out) *visc_out = visc;
    const double EP = visc * W2;
    const double De = 2.0 * visc * visc * A7;
    const double KOL = std::pow(visc * visc * visc / EP, 0.25);
    const double NLAM = (E110 != 0.0) ? 2.0 * A1 / E110 : 0.0;
    const double twopi = 2.0 * double(PI);
    Appender ap{log, logcap};
    if (log && logcap > 0) log[0] = '\0';
    ap.line(" Reynolds n. =%12.1f\n", Re);
    ap.line(" Energy      =%12.7f\n", Q2);
    ap.line(" WiWi        =%12.5f\n", W2);
    ap.line(" Epsilon     =%12.7f\n", EP);
    ap.line(" a11         =%12.7f\n", 2.0 * A1 / Q2 - 1.0);
    ap.line(" e11         =%12.7f\n", 2.0 * (A3 + A4) / W2 - 1.0);
    ap.line(" Time scale  =%12.4f\n", 0.5 * Q2 / EP);
    ap.line(" Kolmogorov  =%12.7f\n", KOL);
    ap.line(" Viscosity   =%12.7f\n", visc);
    ap.line(" dx/Kol.     =%12.7f\n", double(DXZ) / KOL);
    ap.line(" 2Pi/Nlamda  =%12.7f\n", NLAM);
    ap.line(" 2Pi/Lux     =%12.7f\n", twopi / std::sqrt(2.0 * A1 / A3));
    ap.line(" 2Pi/Luz     =%12.7f\n", twopi / std::sqrt(2.0 * A1 / A4));
    ap.line(" 2Pi/Lwx     =%12.7f\n", twopi / std::sqrt(2.0 * A2 / A5));
    ap.line(" 2Pi/Lwz     =%12.7f\n", twopi / std::sqrt(2.0 * A2 / A6));
    ap.line(" Deps.       =%12.7f\n", De);
    ap.line(" Ceps2       =%12.7f\n", 0.5 * Q2 * De / (EP * EP));
    ap.line(" E1          =%12.7f\n", 1.0);
    ap.line(" E3          =%12.7f\n", 1.0);
    ap.iline(" PAO seed    =%12d\n", seed);

    // CALCOM: om2 = i (gamma u1 - alfa u3) at [iz][ix]
    if (om2) {
        for (int z = 0; z < N; ++z) {
            const double gz = double(gam[z]);
            for (int x = 0; x < ND2; ++x) {
                const double ax = double(alfa[x]);
                const f2 a = at(x, z, 0), b = at(x, z, 1);
                const double dre = gz * double(a.re) - ax * double(b.re);
                const double dim = gz * double(a.im) - ax * double(b.im);
                double *o = om2 + 2 * ((size_t)z * ND2 + x);
                o[0] = -dim;
                o[1] = dre;
            }
        }
    }
    return seed;
}

extern "C" dns_status dns_ic_k4exp_host(int N, double K0, uint64_t seed, double *om2)
{
    if (N < 4 || (N % 4) != 0 || !(K0 > 0.0) || !om2) return DNS_E_ARG;
    const int H = N / 2;
    auto gz_of = [N](int iz) { return (iz < N / 2) ? iz : iz - N; };
    // normalization: per-row partial sums added in row order (deterministic)
    std::vector<double> rows(N, 0.0);
    parallel_for(N, [&](int iz) {
        if (iz == N / 2) return;
        const double gz = gz_of(iz);
        double acc = 0.0;
        for (int ix = 0; ix < H; ++ix) {
            if (ix == 0 && iz == 0) continue;
            const double K = std::sqrt(double(ix) * ix + gz * gz);
            acc += ((ix == 0) ? 1.0 : 2.0) * k4exp_mode_energy(K, K0);
        }
        rows[iz] = acc;
    });
    double tot = 0.0;
    for (int iz = 0; iz < N; ++iz) tot += rows[iz];
    const double c = 1.0 / tot;
    parallel_for(N, [&](int iz) {
        const int gz = gz_of(iz);
        for (int ix = 0; ix < H; ++ix) {
            double re = 0.0, im = 0.0;
            if (!(ix == 0 && iz == 0) && iz != N / 2) {
                int kz = gz;
                bool flip = false;
                if (ix == 0 && kz < 0) {
                    kz = -kz;
                    flip = true;
                }
                const double K = std::sqrt(double(ix) * ix + double(kz) * kz);
                const double amp = K * std::sqrt(c * k4exp_mode_energy(K, K0));
                const double th = 2.0 * 3.14159265358979323846 * uniform01(seed, ix, kz);
                re = amp * std::cos(th);
                im = amp * std::sin(th);
                if (flip) im = -im;
            }
            double *o = om2 + 2 * ((size_t)iz * H + ix);
            o[0] = re;
            o[1] = im;
        }
    });
    return DNS_OK;
}
