// dns_ens_kernels.hip -- device side of the batched ensemble: the per-member scalar
// logic of the loop (so a block of steps runs without the host), the snapshot and
// coarse-grid truncation, and the training-pair transform to the n x n grid.
//
// The scalar arithmetic repeats the single-run host code operation by operation, with
// floating-point contraction off, so every member's dt/cn/t sequence -- and with it
// its state -- is bitwise the one of a single run:
//   coefficients  dns_api.cpp dns_step (from step3_cuda.cu:75-100)
//   t, cnm1       mem_cuda.cu:1857, step3_cuda.cu:172
//   NEXTDT        dns_cuda.cu:553-580
#include <hip/hip_runtime.h>

#include "dns_ens_kernels.h"

using dnsfft::cd;

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kPairMaxN = 64;  // coarse grids up to 64 x 64

__global__ void k_ens_pre(EnsScalars *sc, double *coef, int count)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= count) return;
    EnsScalars s = sc[b];
    s.it += 1;
    s.dt_old = s.dt;
    coef[3 * b] = 0.5 * s.visc * s.dt;
    coef[3 * b + 1] = 0.5 * s.dt * (2.0 + s.cnm1);
    coef[3 * b + 2] = -0.5 * s.dt * s.cnm1;
    sc[b] = s;
}

__global__ void k_ens_post(EnsScalars *sc, int count)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= count) return;
    EnsScalars s = sc[b];
    s.t += s.dt_old;
    s.cnm1 = s.cn;
    sc[b] = s;
}

__global__ __launch_bounds__(256) void k_ens_nextdt(EnsScalars *sc, const double *cflrow, int M)
{
#pragma clang fp contract(off)
    __shared__ double red[4];
    const int b = blockIdx.x;
    const double *r = cflrow + (size_t)b * M;
    double m = 0.0;
    for (int i = threadIdx.x; i < M; i += 256) m = r[i] > m ? r[i] : m;
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double cflm = red[0];
    for (int w = 1; w < 4; ++w) cflm = red[w] > cflm ? red[w] : cflm;
    EnsScalars s = sc[b];
    if (s.it == 0 && s.dt == 0.0) {
        s.dt = s.cflnum / cflm / kPi;
        s.cflm = cflm;
    } else if (s.ifn > 0 && (s.it % s.ifn) == 0) {
        const double cfl = cflm * s.dt * kPi;
        s.cn = 0.8 + 0.2 * (s.cflnum / cfl);
        s.dt = s.dt * s.cn;
        s.cflm = cflm;
    }
    sc[b] = s;
}

__global__ void k_ens_copy(const cd *__restrict__ src, cd *__restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// om_c[b][kx][iz] (n-grid, kx < n/2) = om_f[b][kx][kz mod N] for |kz| < n/2, else 0
__global__ void k_ens_truncate(const cd *__restrict__ om_f, int N, cd *__restrict__ om_c, cd *__restrict__ fn_c,
                               int n, int count)
{
    const int hc = n / 2;
    const size_t per = (size_t)n * hc, total = per * count;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t b = e / per;
        const int r = (int)(e - b * per);
        const int kx = r / n, iz = r % n;
        cd v{0.0, 0.0};
        if (iz != hc) {
            const int kz = iz < hc ? iz : iz - n;
            const int izf = kz >= 0 ? kz : kz + N;
            v = om_f[b * (size_t)N * (N / 2) + (size_t)kx * N + izf];
        }
        om_c[e] = v;
        fn_c[e] = cd{0.0, 0.0};
    }
}

// One workgroup per (member, channel): coarse spectrum S[kz][kx] -> physical f[z][x] on
// the n x n grid by two direct DFT passes in LDS (z-inverse, then the x C2R with the
// kx = 0 bin's imaginary part ignored), in FP64; stored as FP32.
__global__ __launch_bounds__(256) void k_ens_pairs(const cd *__restrict__ om_c, const cd *__restrict__ fn_f, int N,
                                                    const cd *__restrict__ fn_c, int n, int count,
                                                    float *__restrict__ pairs, const int *__restrict__ slot,
                                                    int max_pairs)
{
    __shared__ cd S[kPairMaxN * kPairMaxN / 2];
    __shared__ cd A[kPairMaxN * kPairMaxN / 2];
    __shared__ cd E[kPairMaxN];
    const int sl = *slot;
    if (sl >= max_pairs) return;
    const int b = blockIdx.x, ch = blockIdx.y;
    const int hc = n / 2;
    const cd *oc = om_c + (size_t)b * n * hc;
    const cd *fc = fn_c + (size_t)b * n * hc;
    const cd *ff = fn_f + (size_t)b * N * (N / 2);
    for (int m = threadIdx.x; m < n; m += 256) {
        double sv, cv;
        sincospi(2.0 * m / n, &sv, &cv);
        E[m] = cd{cv, sv};
    }
    for (int e = threadIdx.x; e < n * hc; e += 256) {
        const int iz = e / hc, kx = e % hc;
        cd v{0.0, 0.0};
        if (iz != hc) {
            const int kz = iz < hc ? iz : iz - n;
            const cd o = oc[(size_t)kx * n + iz];
            const double ax = double(kx), gz = double(kz);
            const double inv_k2 = 1.0 / (ax * ax + gz * gz + 1.0e-30);
            double s = 0.0;
            switch (ch) {
                case 0: v = o; break;
                case 1:  // u: -i gamma / K^2 (kx = 0: -i / gamma)
                    s = (kx >= 1) ? -gz * inv_k2 : ((iz >= 1 && kz != 0) ? -(1.0 / gz) : 0.0);
                    v = cd{-s * o.y, s * o.x};
                    break;
                case 2:  // w: i alfa / K^2
                    s = (kx >= 1) ? ax * inv_k2 : 0.0;
                    v = cd{-s * o.y, s * o.x};
                    break;
                default: {  // tau = F(FN) - FN of the coarse field
                    const int izf = kz >= 0 ? kz : kz + N;
                    const cd a = ff[(size_t)kx * N + izf], c = fc[(size_t)kx * n + iz];
                    v = cd{a.x - c.x, a.y - c.y};
                    break;
                }
            }
        }
        S[iz * hc + kx] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * hc; e += 256) {
        const int z = e / hc, kx = e % hc;
        double ar = 0.0, ai = 0.0;
        for (int iz = 0; iz < n; ++iz) {
            const cd sv = S[iz * hc + kx], w = E[(iz * z) % n];
            ar += sv.x * w.x - sv.y * w.y;
            ai += sv.x * w.y + sv.y * w.x;
        }
        A[z * hc + kx] = cd{ar, ai};
    }
    __syncthreads();
    float *out = pairs + ((size_t)sl * count + b) * 4 * n * n + (size_t)ch * n * n;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const int z = e / n, x = e % n;
        double f = 0.0;
        for (int kx = 1; kx < hc; ++kx) {
            const cd a = A[z * hc + kx], w = E[(kx * x) % n];
            f += a.x * w.x - a.y * w.y;
        }
        out[e] = float(A[z * hc].x + 2.0 * f);
    }
}

__global__ void k_ens_bump(int *slot)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *slot += 1;
}

int blocks_for(size_t n) { return (int)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384); }

}  // namespace

hipError_t ens_launch_pre(EnsScalars *sc, double *coef, int count, hipStream_t s)
{
    hipLaunchKernelGGL(k_ens_pre, dim3((count + 63) / 64), dim3(64), 0, s, sc, coef, count);
    return hipGetLastError();
}

hipError_t ens_launch_post(EnsScalars *sc, int count, hipStream_t s)
{
    hipLaunchKernelGGL(k_ens_post, dim3((count + 63) / 64), dim3(64), 0, s, sc, count);
    return hipGetLastError();
}

hipError_t ens_launch_nextdt(EnsScalars *sc, const double *cflrow, int M, int count, hipStream_t s)
{
    hipLaunchKernelGGL(k_ens_nextdt, dim3(count), dim3(256), 0, s, sc, cflrow, M);
    return hipGetLastError();
}

hipError_t ens_launch_copy(const cd *src, cd *dst, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_ens_copy, dim3(blocks_for(n)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

hipError_t ens_launch_truncate(const cd *om_f, int N, cd *om_c, cd *fn_c, int n, int count, hipStream_t s)
{
    hipLaunchKernelGGL(k_ens_truncate, dim3(blocks_for((size_t)n * (n / 2) * count)), dim3(256), 0, s, om_f, N,
                       om_c, fn_c, n, count);
    return hipGetLastError();
}

hipError_t ens_launch_pairs(const cd *om_c, const cd *fn_f, int N, const cd *fn_c, int n, int count, float *pairs,
                            const int *slot, int max_pairs, hipStream_t s)
{
    if (n > kPairMaxN || n < 4 || n % 2 != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ens_pairs, dim3(count, 4), dim3(256), 0, s, om_c, fn_f, N, fn_c, n, count, pairs, slot,
                       max_pairs);
    return hipGetLastError();
}

hipError_t ens_launch_bump(int *slot, hipStream_t s)
{
    hipLaunchKernelGGL(k_ens_bump, dim3(1), dim3(64), 0, s, slot);
    return hipGetLastError();
}
