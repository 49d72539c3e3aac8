// cnn_kernels.hip -- the closure network's training and inference kernels for gfx950:
// 3x3 circular convolutions as implicit GEMMs on v_mfma_f32_16x16x32_bf16 (BF16 inputs,
// FP32 accumulation), their data and weight gradients, MSE, Adam and data packing.
//
// Matrix-instruction operand layout (16 x 16 x 32, one wavefront; checked on the device):
//   A (16 x 32): lane l holds row l % 16, k = 8 (l / 16) .. + 7
//   B (32 x 16): lane l holds column l % 16, k = 8 (l / 16) .. + 7
//   D (16 x 16): lane l holds rows 4 (l / 16) .. + 3 of column l % 16
// Forward / data gradient: rows = pixels, columns = output channels, k = (tap, channel)
// so each lane's 8 k-values are 8 adjacent channels of one tap: one 16-byte load.
// Weight gradient: rows = output channels, columns = (tap, channel), k = pixels, staged
// through LDS.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "closure_scaling.hpp"
#include "cnn_kernels.h"

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kp_of(int k) { return (k + 31) / 32 * 32; }

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

__device__ __forceinline__ bf8 load8(const uint16_t *p) { return __builtin_bit_cast(bf8, *reinterpret_cast<const uint4 *>(p)); }

template <int CINP, int COUTP, int EPI>
__global__ __launch_bounds__(256) void k_conv(const uint16_t *__restrict__ in, const uint16_t *__restrict__ W,
                                              const float *__restrict__ bias, int n, int P,
                                              float *__restrict__ outz, uint16_t *__restrict__ outa,
                                              const float *__restrict__ zprev, int cout_real)
{
    constexpr int K = 9 * CINP;
    constexpr int KP = kp_of(K);
    constexpr int NT = COUTP / 16;
    static_assert(CINP % 8 == 0 && COUTP % 16 == 0, "channel padding");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int pix0 = (blockIdx.x * 4 + wave) * 16;
    if (pix0 >= P) return;
    const int nn = n * n;
    const int pa = pix0 + (lane & 15);
    const int b = pa / nn, r = pa - b * nn, y = r / n, x = r - y * n;
    const size_t img = (size_t)b * nn;
    const int kq = 8 * (lane >> 4);
    f4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 3
    for (int ks = 0; ks < KP; ks += 32) {
        const int k = ks + kq;
        bf8 a = __builtin_bit_cast(bf8, uint4{0u, 0u, 0u, 0u});
        if (k < K) {
            const int tap = k / CINP, ci = k - tap * CINP;
            int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
            yy += (yy < 0) ? n : ((yy >= n) ? -n : 0);
            xx += (xx < 0) ? n : ((xx >= n) ? -n : 0);
            a = load8(in + (img + (size_t)yy * n + xx) * CINP + ci);
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf8 w = load8(W + (size_t)(t * 16 + (lane & 15)) * KP + k);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t p = (size_t)pix0 + 4 * (lane >> 4) + j;
            float v = acc[t][j];
            if (EPI == CONV_BIAS_GELU) {
                v += bias[co];
                outz[p * COUTP + co] = v;
                outa[p * COUTP + co] = f2bf(gelu(v));
            } else if (EPI == CONV_BIAS) {
                if (co < cout_real) outz[p * cout_real + co] = v + bias[co];
            } else {
                outa[p * COUTP + co] = f2bf(v * gelu_grad(zprev[p * COUTP + co]));
            }
        }
    }
}

template <int CINP, int COUTP>
__global__ __launch_bounds__(256) void k_wgrad(const uint16_t *__restrict__ dz, const uint16_t *__restrict__ in, int n,
                                               int P, int ppb, float *__restrict__ part, float *__restrict__ partb)
{
    constexpr int K = 9 * CINP;
    constexpr int NTK = (K + 15) / 16;
    constexpr int MT = COUTP / 16;
    constexpr int TILES = MT * NTK;
    constexpr int PER = (TILES + 3) / 4;
    __shared__ uint16_t sdz[32 * COUTP];
    __shared__ uint16_t sin[9 * 32 * CINP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nn = n * n;
    const int p0 = blockIdx.x * ppb;
    const int p1 = min(P, p0 + ppb);
    f4 acc[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) acc[i] = f4{0.f, 0.f, 0.f, 0.f};
    float sumb = 0.0f;
    for (int q0 = p0; q0 < p1; q0 += 32) {
        // stage dz[32][COUTP] and the 9 shifted input rows [tap][32][CINP] (8 channels a load)
        for (int e = threadIdx.x; e < 32 * COUTP / 8; e += 256) {
            const int px = e / (COUTP / 8), c8 = e % (COUTP / 8);
            *reinterpret_cast<uint4 *>(&sdz[px * COUTP + 8 * c8]) =
                *reinterpret_cast<const uint4 *>(dz + (size_t)(q0 + px) * COUTP + 8 * c8);
        }
        for (int e = threadIdx.x; e < 9 * 32 * CINP / 8; e += 256) {
            const int c8 = e % (CINP / 8), rest = e / (CINP / 8);
            const int px = rest % 32, tap = rest / 32;
            const int p = q0 + px;
            const int b = p / nn, r = p - b * nn, y = r / n, x = r - y * n;
            int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
            yy += (yy < 0) ? n : ((yy >= n) ? -n : 0);
            xx += (xx < 0) ? n : ((xx >= n) ? -n : 0);
            *reinterpret_cast<uint4 *>(&sin[(tap * 32 + px) * CINP + 8 * c8]) =
                *reinterpret_cast<const uint4 *>(in + ((size_t)b * nn + (size_t)yy * n + xx) * CINP + 8 * c8);
        }
        __syncthreads();
        if (threadIdx.x < COUTP)
            for (int px = 0; px < 32; ++px) sumb += __uint_as_float((uint32_t)sdz[px * COUTP + threadIdx.x] << 16);
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int tile = wave + 4 * i;
            if (tile < TILES) {
                const int mt = tile / NTK, nt = tile % NTK;
                uint16_t av[8], bv[8];
                const int co = mt * 16 + (lane & 15);
                const int k = nt * 16 + (lane & 15);
                const int tap = k / CINP, ci = k - tap * CINP;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int px = 8 * (lane >> 4) + j;
                    av[j] = sdz[px * COUTP + co];
                    bv[j] = (k < K) ? sin[(tap * 32 + px) * CINP + ci] : (uint16_t)0;
                }
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, av), __builtin_bit_cast(bf8, bv),
                                                                 acc[i], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    float *out = part + (size_t)blockIdx.x * COUTP * K;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int tile = wave + 4 * i;
        if (tile < TILES) {
            const int mt = tile / NTK, nt = tile % NTK;
            const int k = nt * 16 + (lane & 15);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int co = mt * 16 + 4 * (lane >> 4) + j;
                if (k < K) out[(size_t)co * K + k] = acc[i][j];
            }
        }
    }
    if (threadIdx.x < COUTP) partb[(size_t)blockIdx.x * COUTP + threadIdx.x] = sumb;
}

__global__ void k_wgrad_reduce(int cinp, int coutp, int cin, int cout, const float *__restrict__ part,
                               const float *__restrict__ partb, int blocks, float *__restrict__ grad, size_t off)
{
    const int K = 9 * cinp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw = cout * cin * 9;
    if (i < nw) {
        // flat order W[co][ci][ky][kx]
        const int co = i / (cin * 9), rem = i - co * cin * 9, ci = rem / 9, tap = rem - ci * 9;
        const int k = tap * cinp + ci;
        float s = 0.0f;
        for (int b = 0; b < blocks; ++b) s += part[((size_t)b * coutp + co) * K + k];
        grad[off + i] = s;
    } else if (i < nw + cout) {
        const int co = i - nw;
        float s = 0.0f;
        for (int b = 0; b < blocks; ++b) s += partb[(size_t)b * coutp + co];
        grad[off + i] = s;
    }
}

__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ y, const float *__restrict__ t, int P,
                                             double *__restrict__ lossp, uint16_t *__restrict__ dz)
{
    __shared__ double red[4];
    double s = 0.0;
    const float scale = 2.0f / float(P);
    for (int p = blockIdx.x * 256 + threadIdx.x; p < P; p += gridDim.x * 256) {
        const float d = y[p] - t[p];
        s += double(d) * double(d);
        uint4 z0{0u, 0u, 0u, 0u};
        z0.x = f2bf(scale * d);
        *reinterpret_cast<uint4 *>(dz + (size_t)p * 16) = z0;
        *reinterpret_cast<uint4 *>(dz + (size_t)p * 16 + 8) = uint4{0u, 0u, 0u, 0u};
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) lossp[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       int count, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void k_pack_layer(const float *__restrict__ p, size_t off, int cin, int cout, int cinp, int coutp,
                             uint16_t *__restrict__ Wf, uint16_t *__restrict__ Wb, float *__restrict__ bias)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw = cout * cin * 9;
    const int kpf = kp_of(9 * cinp), kpb = kp_of(9 * coutp);
    if (i < nw) {
        const int co = i / (cin * 9), rem = i - co * cin * 9, ci = rem / 9, tap = rem - ci * 9;
        const uint16_t w = f2bf(p[off + i]);
        Wf[(size_t)co * kpf + tap * cinp + ci] = w;
        if (Wb) Wb[(size_t)ci * kpb + (8 - tap) * coutp + co] = w;
    } else if (i < nw + cout) {
        bias[i - nw] = p[off + i];
    }
}

__global__ void k_pack_input(const float *__restrict__ x, int B, int n, uint16_t *__restrict__ a0)
{
    const int nn = n * n;
    const size_t P = (size_t)B * nn;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
        const size_t b = p / nn, r = p - b * nn;
        const float *xb = x + b * 3 * nn + r;
        uint4 v{0u, 0u, 0u, 0u};
        v.x = (uint32_t)f2bf(xb[0]) | ((uint32_t)f2bf(xb[nn]) << 16);
        v.y = (uint32_t)f2bf(xb[2 * (size_t)nn]);
        *reinterpret_cast<uint4 *>(a0 + p * 8) = v;
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_gather_pairs(const float *__restrict__ pairs, int count, int stored, int n,
                                                      uint64_t seed, int iter, int B, uint16_t *__restrict__ a0,
                                                      float *__restrict__ t)
{
    using closure_scaling::highpass;
    __shared__ double red[4];
    const int s = blockIdx.x;
    const uint64_t h = mix64(seed ^ mix64((uint64_t)iter * (uint64_t)B + (uint64_t)s));
    const uint64_t idx = h % (uint64_t)((uint64_t)stored * count);
    const int nn = n * n;
    const float *src = pairs + idx * 4 * (size_t)nn;  // pair (slot, member) = (idx / count, idx % count)
    double w2 = 0.0;
    for (int p = threadIdx.x; p < nn; p += 256) {
        const double w = highpass(src, n, p / n, p % n);
        w2 += w * w;
    }
    for (int o = 32; o > 0; o >>= 1) w2 += __shfl_xor(w2, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w2;
    __syncthreads();
    const double ws = sqrt((red[0] + red[1] + red[2] + red[3]) / nn);
    const double D = 6.283185307179586476925 / n;
    const float iw = ws > 1e-30 ? float(1.0 / ws) : 0.0f;
    const float iu = ws > 1e-30 ? float(1.0 / (ws * D)) : 0.0f;
    const float it = ws > 1e-30 ? float(1.0 / (ws * ws)) : 0.0f;
    for (int p = threadIdx.x; p < nn; p += 256) {
        const int y = p / n, x = p % n;
        uint4 v{0u, 0u, 0u, 0u};
        v.x = (uint32_t)f2bf(highpass(src, n, y, x) * iw) | ((uint32_t)f2bf(highpass(src + nn, n, y, x) * iu) << 16);
        v.y = (uint32_t)f2bf(highpass(src + 2 * (size_t)nn, n, y, x) * iu);
        *reinterpret_cast<uint4 *>(a0 + ((size_t)s * nn + p) * 8) = v;
        t[(size_t)s * nn + p] = src[3 * nn + p] * it;
    }
}

}  // namespace

hipError_t cnn_launch_conv(int cinp, int coutp, int epi, const uint16_t *in, const uint16_t *W, const float *bias,
                           int n, int P, float *outz, uint16_t *outa, const float *zprev, int cout_real, hipStream_t s)
{
    if (P % 16 != 0) return hipErrorInvalidValue;
    const dim3 grid((P / 16 + 3) / 4), block(256);
#define CONV_CASE(CI, CO, E)                                                                                  \
    if (cinp == CI && coutp == CO && epi == E) {                                                              \
        hipLaunchKernelGGL((k_conv<CI, CO, E>), grid, block, 0, s, in, W, bias, n, P, outz, outa, zprev,     \
                           cout_real);                                                                        \
        return hipGetLastError();                                                                             \
    }
    CONV_CASE(8, 32, CONV_BIAS_GELU)
    CONV_CASE(32, 32, CONV_BIAS_GELU)
    CONV_CASE(32, 16, CONV_BIAS)
    CONV_CASE(16, 32, CONV_GELU_GRAD)
    CONV_CASE(32, 32, CONV_GELU_GRAD)
#undef CONV_CASE
    return hipErrorInvalidValue;
}

hipError_t cnn_launch_wgrad(int cinp, int coutp, const uint16_t *dz, const uint16_t *in, int n, int P, int ppb,
                            float *part, float *partb, hipStream_t s)
{
    if (P % 32 != 0 || ppb % 32 != 0) return hipErrorInvalidValue;
    const dim3 grid((P + ppb - 1) / ppb), block(256);
#define WG_CASE(CI, CO)                                                                                   \
    if (cinp == CI && coutp == CO) {                                                                      \
        hipLaunchKernelGGL((k_wgrad<CI, CO>), grid, block, 0, s, dz, in, n, P, ppb, part, partb);         \
        return hipGetLastError();                                                                         \
    }
    WG_CASE(8, 32)
    WG_CASE(32, 32)
    WG_CASE(32, 16)
#undef WG_CASE
    return hipErrorInvalidValue;
}

hipError_t cnn_launch_wgrad_reduce(int cinp, int coutp, int cin, int cout, const float *part, const float *partb,
                                   int blocks, float *grad, size_t off, hipStream_t s)
{
    const int total = cout * cin * 9 + cout;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((total + 255) / 256), dim3(256), 0, s, cinp, coutp, cin, cout, part, partb,
                       blocks, grad, off);
    return hipGetLastError();
}

hipError_t cnn_launch_mse(const float *y, const float *t, int P, double *lossp, int blocks, uint16_t *dz, hipStream_t s)
{
    hipLaunchKernelGGL(k_mse, dim3(blocks), dim3(256), 0, s, y, t, P, lossp, dz);
    return hipGetLastError();
}

hipError_t cnn_launch_adam(float *p, const float *g, float *m, float *v, int count, float lr, float b1, float b2,
                           float eps, float c1, float c2, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3((count + 255) / 256), dim3(256), 0, s, p, g, m, v, count, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t cnn_launch_pack_layer(const float *p, size_t off, int cin, int cout, int cinp, int coutp, uint16_t *Wf,
                                 uint16_t *Wb, float *bias, hipStream_t s)
{
    const int total = cout * cin * 9 + cout;
    hipLaunchKernelGGL(k_pack_layer, dim3((total + 255) / 256), dim3(256), 0, s, p, off, cin, cout, cinp, coutp, Wf, Wb,
                       bias);
    return hipGetLastError();
}

hipError_t cnn_launch_pack_input(const float *x, int B, int n, uint16_t *a0, hipStream_t s)
{
    const size_t P = (size_t)B * n * n;
    const int blocks = (int)std::min<size_t>((P + 255) / 256, 16384);
    hipLaunchKernelGGL(k_pack_input, dim3(blocks), dim3(256), 0, s, x, B, n, a0);
    return hipGetLastError();
}

hipError_t cnn_launch_gather_pairs(const float *pairs, int count, int stored, int n, int B, uint64_t seed, int iter,
                                   uint16_t *a0, float *t, hipStream_t s)
{
    if (stored <= 0 || count <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gather_pairs, dim3(B), dim3(256), 0, s, pairs, count, stored, n, seed, iter, B, a0, t);
    return hipGetLastError();
}
