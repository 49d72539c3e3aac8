// dns_big.hip -- the time step for grids whose half-width H = N/2 is too long for the
// fused LDS-resident row/column kernels (e.g. N = 8192, N = 29400: 44100^2 padded).
//
// The state (om, fn: [kx][iz]) is the same as in the fused path; the transforms run
// as batched M-point complex line FFTs through HBM, each split M = L1 x L2 into two
// LDS-resident passes (four-step):
//   pass A  (per line, C adjacent n2):  y[k1 L2 + n2] = W_M^{s n2 k1} FFT_L1(x[n1 L2 + n2])[k1]
//   pass B  (per line, C adjacent k1):  X[k1 + L1 k2] = FFT_L2(y[k1 L2 + n2])[k2]
// One step (reference order mem_cuda.cu:1851-1857; FN in advective form):
//   spec    om -> 4 padded spectra (u, w, d_x omega, d_z omega) as z-lines  [f][kx][z]
//   z-inv   4H lines of M
//   pack    (u, w) and (d_x omega, d_z omega) -> full x-spectra a + i b of each row
//           (Hermitian extension; LDS-tiled transpose)                     [p][z][x]
//   x-inv   2M lines of M  -> (u + i w), (w_x + i w_z) on the M x M grid
//   prod    P = u w_x + w w_z, two rows packed as one complex line          [M/2][x]
//   x-fwd   M/2 lines of M
//   unpack  split the row pair, keep kx < H, transpose to z-lines          [kx][z]
//   z-fwd   H lines of M
//   cn      FN = -P^/M^2 (kz = -N/2 row 0), CN/AB2 update of om, fn
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "dns_big.h"
#include "fft_lds.hpp"

using namespace dnsfft;

namespace {

__device__ __forceinline__ double gam(int iz, int N) { return (iz < (N >> 1)) ? double(iz) : double(iz - N); }

// ---------------------------------------------------------------------------------
// batched four-step line FFT, pass A
// ---------------------------------------------------------------------------------
template <class TC, int L1, int L2, int C, int T, bool INV>
__global__ __launch_bounds__(T) void k_big_fft_a(const TC *__restrict__ x, TC *__restrict__ y,
                                                 const TC *__restrict__ tw)
{
    constexpr int L = L1 * L2;
    constexpr int LS = padded_len(L1);
    __shared__ TC lds[C * LS];
    const size_t line = blockIdx.y;
    const int n20 = blockIdx.x * C;
    const TC *src = x + line * L;
    TC *dst = y + line * L;
    for (int e = threadIdx.x; e < C * L1; e += T) {
        const int c = e % C, n1 = e / C;
        lds[c * LS + pidx(n1)] = src[(size_t)n1 * L2 + n20 + c];
    }
    __syncthreads();
    fft_lines<L1, LS, C, T, INV>(lds, tw, L2);
    for (int e = threadIdx.x; e < C * L1; e += T) {
        const int c = e % C, k1 = e / C;
        const int n2 = n20 + c;
        const TC w = tw[opaque((int)(((long long)n2 * k1) % L))];
        const TC v = lds[c * LS + pidx(k1)];
        dst[(size_t)k1 * L2 + n2] = INV ? mulc(v, w) : mul(v, w);
    }
}

// ---------------------------------------------------------------------------------
// batched four-step line FFT, pass B (natural-order output)
// ---------------------------------------------------------------------------------
template <class TC, int L1, int L2, int C, int T, bool INV>
__global__ __launch_bounds__(T) void k_big_fft_b(const TC *__restrict__ y, TC *__restrict__ out,
                                                 const TC *__restrict__ tw)
{
    constexpr int L = L1 * L2;
    constexpr int LS = padded_len(L2);
    __shared__ TC lds[C * LS];
    const size_t line = blockIdx.y;
    const int k10 = blockIdx.x * C;
    const TC *src = y + line * L;
    TC *dst = out + line * L;
    for (int e = threadIdx.x; e < C * L2; e += T) {
        const int c = e / L2, n2 = e % L2;
        lds[c * LS + pidx(n2)] = src[(size_t)(k10 + c) * L2 + n2];
    }
    __syncthreads();
    fft_lines<L2, LS, C, T, INV>(lds, tw, L1);
    for (int e = threadIdx.x; e < C * L2; e += T) {
        const int c = e % C, k2 = e / C;
        dst[(size_t)k2 * L1 + k10 + c] = lds[c * LS + pidx(k2)];
    }
}

// spectrum multiplier of field f (the fused column kernel's conventions: the
// multipliers of omega listed with DNS_F_*)
__device__ __forceinline__ cd spec_of(int f, cd om, int ix, int iz, int N)
{
    const double ax = double(ix), gz = gam(iz, N);
    const double inv_k2 = 1.0 / (ax * ax + gz * gz + 1.0e-30);
    double s;
    bool ti = true;
    switch (f) {
        case DNS_F_U: s = (ix >= 1) ? -gz * inv_k2 : ((iz >= 1 && gz != 0.0) ? -(1.0 / gz) : 0.0); break;
        case DNS_F_W: s = (ix >= 1) ? ax * inv_k2 : 0.0; break;
        case DNS_F_OX: s = ax; break;
        case DNS_F_OZ: s = gz; break;
        case DNS_F_PSI:
            s = (ix == 0 && iz == 0) ? 0.0 : inv_k2;
            ti = false;
            break;
        default:
            s = 1.0;
            ti = false;
            break;
    }
    return ti ? cd{-s * om.y, s * om.x} : cd{s * om.x, s * om.y};
}

// om [kx][iz] -> padded z-lines Z[fi][kx][z'] (z' in [0, M); the band [H, 2H) is zero)
template <class TC>
__global__ void k_big_spec(const cd *__restrict__ om, TC *__restrict__ Z, BigSpecArgs a, int H)
{
    const int N = 2 * H, M = 3 * H;
    const size_t total = (size_t)H * M;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(e / M), zp = (int)(e % M);
        int iz = -1;
        if (zp < H)
            iz = zp;
        else if (zp >= 2 * H)
            iz = zp - H;
        const cd o = (iz >= 0) ? om[(size_t)kx * N + iz] : cd{0.0, 0.0};
        for (int fi = 0; fi < a.nfields; ++fi) {
            const cd v = (iz >= 0) ? spec_of(a.fields[fi], o, kx, iz, N) : cd{0.0, 0.0};
            Z[(size_t)fi * total + e] = from_d<TC>(v);
        }
    }
}

// Row x-spectra of two fields a, b (z-lines, [kx][z]) packed as a + i b with the
// Hermitian extension, transposed through a 32 x 32 LDS tile into rows R[z][x].
template <class TC>
__global__ __launch_bounds__(256) void k_big_pack(const TC *__restrict__ Za, const TC *__restrict__ Zb,
                                                  TC *__restrict__ R, int H)
{
    using RT = decltype(TC::x);
    const int M = 3 * H;
    __shared__ TC lo[32][33];
    __shared__ TC hi[32][33];
    const int kx0 = blockIdx.x * 32, z0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int kx = kx0 + r, z = z0 + tx;
        TC zl{RT(0), RT(0)}, zh{RT(0), RT(0)};
        if (kx < H && z < M) {
            const TC av = Za[(size_t)kx * M + z];
            const TC bv = Zb ? Zb[(size_t)kx * M + z] : TC{RT(0), RT(0)};
            if (kx == 0) {
                zl = TC{av.x, bv.x};
            } else {
                zl = TC{av.x - bv.y, av.y + bv.x};
                zh = TC{av.x + bv.y, bv.x - av.y};  // conj(a) + i conj(b)
            }
        }
        lo[r][tx] = zl;
        hi[r][tx] = zh;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int z = z0 + r, kx = kx0 + tx;
        if (z < M && kx < H) {
            TC *row = R + (size_t)z * M;
            row[kx] = lo[tx][r];
            if (kx >= 1) row[M - kx] = hi[tx][r];
        }
    }
}

// zero the band kx in [H, M - H] of every row (the mirror fills (M - H, M))
template <class TC>
__global__ void k_big_zero_band(TC *__restrict__ R, int H, int rows)
{
    using RT = decltype(TC::x);
    const int M = 3 * H, w = M - 2 * H + 1;
    const size_t total = (size_t)rows * w;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t z = e / w;
        R[z * M + H + (e % w)] = TC{RT(0), RT(0)};
    }
}

// P = u w_x + w w_z from RA = u + i w, RB = w_x + i w_z; rows 2q, 2q+1 packed into Q[q];
// optional per-row-pair CFL maxima
template <class TC>
__global__ __launch_bounds__(256) void k_big_prod(const TC *__restrict__ RA, const TC *__restrict__ RB,
                                                  TC *__restrict__ Q, int M, double *__restrict__ cflrow,
                                                  double DX, double DZ)
{
    __shared__ double red[4];
    const int q = blockIdx.x;
    double cm = 0.0;
    for (int x = threadIdx.x; x < M; x += 256) {
        const TC a0 = RA[(size_t)(2 * q) * M + x], b0 = RB[(size_t)(2 * q) * M + x];
        const TC a1 = RA[(size_t)(2 * q + 1) * M + x], b1 = RB[(size_t)(2 * q + 1) * M + x];
        Q[(size_t)q * M + x] = TC{a0.x * b0.x + a0.y * b0.y, a1.x * b1.x + a1.y * b1.y};
        const double c0 = fabs(double(a0.x)) / DX + fabs(double(a0.y)) / DZ;
        const double c1 = fabs(double(a1.x)) / DX + fabs(double(a1.y)) / DZ;
        cm = fmax(cm, fmax(c0, c1));
    }
    if (cflrow) {
        for (int off = 32; off > 0; off >>= 1) cm = fmax(cm, __shfl_xor(cm, off, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cm;
        __syncthreads();
        if (threadIdx.x == 0) cflrow[q] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    }
}

// per-row CFL maxima of R = u + i w
template <class TC>
__global__ __launch_bounds__(256) void k_big_cfl(const TC *__restrict__ R, int M, double *__restrict__ cflrow,
                                                 double DX, double DZ)
{
    __shared__ double red[4];
    const int z = blockIdx.x;
    double cm = 0.0;
    for (int x = threadIdx.x; x < M; x += 256) {
        const TC v = R[(size_t)z * M + x];
        cm = fmax(cm, fabs(double(v.x)) / DX + fabs(double(v.y)) / DZ);
    }
    for (int off = 32; off > 0; off >>= 1) cm = fmax(cm, __shfl_xor(cm, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cm;
    __syncthreads();
    if (threadIdx.x == 0) cflrow[z] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Q^ (row pairs) -> the two rows' forward spectra for kx < H, transposed into z-lines
// G[kx][z]: F_a = (Z[k] + conj Z[M-k]) / 2, F_b = (Z[k] - conj Z[M-k]) / 2i
template <class TC>
__global__ __launch_bounds__(256) void k_big_unpack(const TC *__restrict__ Q, TC *__restrict__ G, int H)
{
    using RT = decltype(TC::x);
    const int M = 3 * H;
    __shared__ TC ta[32][33];
    __shared__ TC tb[32][33];
    const int kx0 = blockIdx.x * 32, q0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const RT half = RT(0.5);
    for (int r = ty; r < 32; r += 8) {
        const int q = q0 + r, kx = kx0 + tx;
        TC fa{RT(0), RT(0)}, fb{RT(0), RT(0)};
        if (q < M / 2 && kx < H) {
            const TC zk = Q[(size_t)q * M + kx];
            const TC zm = conj(Q[(size_t)q * M + (kx == 0 ? 0 : M - kx)]);
            const TC s = add(zk, zm), d = sub(zk, zm);
            fa = TC{half * s.x, half * s.y};
            fb = TC{half * d.y, -half * d.x};
        }
        ta[r][tx] = fa;
        tb[r][tx] = fb;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int kx = kx0 + r, q = q0 + tx;
        if (kx < H && q < M / 2) {
            G[(size_t)kx * M + 2 * q] = ta[tx][r];
            G[(size_t)kx * M + 2 * q + 1] = tb[tx][r];
        }
    }
}

// FN from the z-forward lines and the CN/AB2 update (step3_cuda.cu:75-100)
template <class TC>
__global__ void k_big_cn(cd *__restrict__ om, cd *__restrict__ fn, const TC *__restrict__ G, BigCnArgs a, int H)
{
    const int N = 2 * H, M = 3 * H;
    const size_t total = (size_t)H * N;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(e / N), iz = (int)(e % N);
        const int zp = (iz < H) ? iz : iz + H;
        const cd X = to_d(G[(size_t)kx * M + zp]);
        cd FN{-X.x * a.divxz, -X.y * a.divxz};
        if (iz == H) FN = cd{0.0, 0.0};
        const double ax = double(kx), gz = gam(iz, N);
        const double ARG = a.VT * (ax * ax + gz * gz);
        const double DEN = 1.0 + ARG, c1 = 1.0 - ARG;
        const cd o = om[e], f = fn[e];
        const cd num{c1 * o.x + a.c2 * FN.x + a.c3 * f.x, c1 * o.y + a.c2 * FN.y + a.c3 * f.y};
        om[e] = cd{num.x / DEN, num.y / DEN};
        fn[e] = FN;
    }
}

// physical planes from R = a + i b (FP64): out0 = Re, out1 = Im (optional) or |R|
template <class TC>
__global__ void k_big_phys(const TC *__restrict__ R, size_t n, double *__restrict__ out0, double *__restrict__ out1,
                           int out_mode)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const cd v = to_d(R[e]);
        if (out_mode == 1) {
            out0[e] = sqrt(v.x * v.x + v.y * v.y);
        } else {
            out0[e] = v.x;
            if (out1) out1[e] = v.y;
        }
    }
}

// ---------------------------------------------------------------------------------
// line-length table: M = L1 x L2, C adjacent sub-lines per workgroup (C divides both)
// ---------------------------------------------------------------------------------
template <int M_>
struct BigLen;
#define BIG_LEN(M_, L1_, L2_, C_)          \
    template <>                            \
    struct BigLen<M_> {                    \
        static constexpr int L1 = L1_;     \
        static constexpr int L2 = L2_;     \
        static constexpr int C = C_;       \
    };
BIG_LEN(96, 8, 12, 4)
BIG_LEN(192, 12, 16, 4)
BIG_LEN(384, 16, 24, 8)
BIG_LEN(768, 24, 32, 8)
BIG_LEN(6144, 64, 96, 16)
BIG_LEN(1536, 32, 48, 16)
BIG_LEN(12288, 96, 128, 16)
BIG_LEN(24576, 128, 192, 16)
BIG_LEN(44100, 210, 210, 14)
#undef BIG_LEN

template <class TC, int M_, bool INV>
hipError_t line_fft_len(const TC *x, TC *y, TC *out, const TC *tw, size_t lines, hipStream_t s)
{
    using B = BigLen<M_>;
    static_assert(B::L1 * B::L2 == M_ && B::L1 % B::C == 0 && B::L2 % B::C == 0, "bad four-step split");
    constexpr int T = 256;
    hipLaunchKernelGGL((k_big_fft_a<TC, B::L1, B::L2, B::C, T, INV>), dim3(B::L2 / B::C, (unsigned)lines), dim3(T),
                       0, s, x, y, tw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_big_fft_b<TC, B::L1, B::L2, B::C, T, INV>), dim3(B::L1 / B::C, (unsigned)lines), dim3(T),
                       0, s, y, out, tw);
    return hipGetLastError();
}

template <class TC, bool INV>
hipError_t line_fft(int M, const TC *x, TC *y, TC *out, const TC *tw, size_t lines, hipStream_t s)
{
    switch (M) {
        case 96: return line_fft_len<TC, 96, INV>(x, y, out, tw, lines, s);
        case 192: return line_fft_len<TC, 192, INV>(x, y, out, tw, lines, s);
        case 384: return line_fft_len<TC, 384, INV>(x, y, out, tw, lines, s);
        case 768: return line_fft_len<TC, 768, INV>(x, y, out, tw, lines, s);
        case 1536: return line_fft_len<TC, 1536, INV>(x, y, out, tw, lines, s);
        case 6144: return line_fft_len<TC, 6144, INV>(x, y, out, tw, lines, s);
        case 12288: return line_fft_len<TC, 12288, INV>(x, y, out, tw, lines, s);
        case 24576: return line_fft_len<TC, 24576, INV>(x, y, out, tw, lines, s);
        case 44100: return line_fft_len<TC, 44100, INV>(x, y, out, tw, lines, s);
        default: return hipErrorInvalidValue;
    }
}

int grid_for(size_t n) { return (int)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536); }

template <class TC>
hipError_t launch_big_t(const BigReq &r)
{
    const int H = r.H, M = 3 * H;
    const TC *tw = static_cast<const TC *>(r.tw);
    hipStream_t s = r.stream;
    switch (r.what) {
        case BIG_SPEC:
            hipLaunchKernelGGL(k_big_spec<TC>, dim3(grid_for((size_t)H * M)), dim3(256), 0, s, r.om,
                               static_cast<TC *>(r.a), r.spec, H);
            return hipGetLastError();
        case BIG_FFT_INV:
            return line_fft<TC, true>(M, static_cast<const TC *>(r.a), static_cast<TC *>(r.scratch),
                                      static_cast<TC *>(r.b), tw, r.lines, s);
        case BIG_FFT_FWD:
            return line_fft<TC, false>(M, static_cast<const TC *>(r.a), static_cast<TC *>(r.scratch),
                                       static_cast<TC *>(r.b), tw, r.lines, s);
        case BIG_PACK: {
            hipLaunchKernelGGL(k_big_zero_band<TC>, dim3(grid_for((size_t)M * (H + 1))), dim3(256), 0, s,
                               static_cast<TC *>(r.b), H, M);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_big_pack<TC>, dim3((H + 31) / 32, (M + 31) / 32), dim3(256), 0, s,
                               static_cast<const TC *>(r.a), static_cast<const TC *>(r.a2), static_cast<TC *>(r.b),
                               H);
            return hipGetLastError();
        }
        case BIG_PROD:
            hipLaunchKernelGGL(k_big_prod<TC>, dim3(M / 2), dim3(256), 0, s, static_cast<const TC *>(r.a),
                               static_cast<const TC *>(r.a2), static_cast<TC *>(r.b), M, r.cflrow, r.DX, r.DZ);
            return hipGetLastError();
        case BIG_CFL:
            hipLaunchKernelGGL(k_big_cfl<TC>, dim3(M), dim3(256), 0, s, static_cast<const TC *>(r.a), M, r.cflrow,
                               r.DX, r.DZ);
            return hipGetLastError();
        case BIG_UNPACK:
            hipLaunchKernelGGL(k_big_unpack<TC>, dim3((H + 31) / 32, (M / 2 + 31) / 32), dim3(256), 0, s,
                               static_cast<const TC *>(r.a), static_cast<TC *>(r.b), H);
            return hipGetLastError();
        case BIG_CN:
            hipLaunchKernelGGL(k_big_cn<TC>, dim3(grid_for((size_t)H * 2 * H)), dim3(256), 0, s, r.om, r.fn,
                               static_cast<const TC *>(r.a), r.cn, H);
            return hipGetLastError();
        case BIG_PHYS:
            hipLaunchKernelGGL(k_big_phys<TC>, dim3(grid_for((size_t)M * M)), dim3(256), 0, s,
                               static_cast<const TC *>(r.a), (size_t)M * M, r.out0, r.out1, r.out_mode);
            return hipGetLastError();
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace

int dns_big_supported_m(int M)
{
    switch (M) {
        case 96: case 192: case 384: case 768: case 1536: case 6144: case 12288: case 24576: case 44100:
            return 1;
        default: return 0;
    }
}

hipError_t dns_big_launch(const BigReq &r)
{
    return r.fp32 ? launch_big_t<cf>(r) : launch_big_t<cd>(r);
}
