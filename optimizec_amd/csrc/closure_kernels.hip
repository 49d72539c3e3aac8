// closure_kernels.hip -- transforms between a coarse run's spectral state and the N x N
// physical grid the closure network works on (no 3/2 padding: the network sees the
// fields at x_j = 2 pi j / N, like the ensemble's training pairs).
//   to grid    omega, u, w: z-inverse lines (N points, 4 kx columns a workgroup), then
//              the x C2R of (omega + i u) and (w + i 0) per row, FP64; FP32 fields out
//   pack       small-scale fields and their vorticity rms w_s (closure_scaling.hpp),
//              bf16 channels-last network input
//   from grid  network output times w_s^2: x-forward of row pairs packed as one
//              complex line, split into the two rows' spectra (kx < N/2), then
//              z-forward lines, / N^2, kz = -N/2 row 0 -> forcing [kx][iz]
#include <hip/hip_runtime.h>

#include "closure_internal.h"
#include "closure_scaling.hpp"
#include "fft_lds.hpp"

using namespace dnsfft;

namespace {

constexpr int kNL = 4;  // z-lines per workgroup

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// multipliers of omega-hat: 0 omega, 1 u = -i gamma / K^2 (kx = 0: -i / gamma), 2 w = i alfa / K^2
__device__ __forceinline__ cd field_of(int f, cd o, int kx, int iz, int N)
{
    if (f == 0) return o;
    const double ax = double(kx), gz = (iz < N / 2) ? double(iz) : double(iz - N);
    const double inv_k2 = 1.0 / (ax * ax + gz * gz + 1.0e-30);
    double s;
    if (f == 1)
        s = (kx >= 1) ? -gz * inv_k2 : ((iz >= 1 && gz != 0.0) ? -(1.0 / gz) : 0.0);
    else
        s = (kx >= 1) ? ax * inv_k2 : 0.0;
    return cd{-s * o.y, s * o.x};
}

template <int N, int T>
__global__ __launch_bounds__(T) void k_to_grid_z(const cd *__restrict__ om, const cd *__restrict__ twN,
                                                 cd *__restrict__ az)
{
    constexpr int H = N / 2, LS = padded_len(N);
    __shared__ cd lds[kNL * LS];
    const int f = blockIdx.y, kx0 = blockIdx.x * kNL;
    for (int e = threadIdx.x; e < kNL * N; e += T) {
        const int c = e / N, iz = e - c * N;
        const int kx = kx0 + c;
        lds[c * LS + pidx(iz)] = (iz == H) ? cd{0.0, 0.0} : field_of(f, om[(size_t)kx * N + iz], kx, iz, N);
    }
    __syncthreads();
    fft_lines<N, LS, kNL, T, true>(lds, twN, 1);
    for (int e = threadIdx.x; e < kNL * N; e += T) {
        const int c = e % kNL, z = e / kNL;
        az[((size_t)f * N + z) * H + kx0 + c] = lds[c * LS + pidx(z)];
    }
}

template <int N, int T>
__global__ __launch_bounds__(T) void k_to_grid_x(const cd *__restrict__ az, const cd *__restrict__ twN,
                                                 float *__restrict__ fields)
{
    constexpr int H = N / 2, LS = padded_len(N);
    __shared__ cd lds[2 * LS];
    const int z = blockIdx.x;
    const cd *aw = az + ((size_t)0 * N + z) * H;
    const cd *au = az + ((size_t)1 * N + z) * H;
    const cd *av = az + ((size_t)2 * N + z) * H;
    for (int k = threadIdx.x; k < N; k += T) {
        cd l1{0.0, 0.0}, l2{0.0, 0.0};
        if (k == 0) {
            l1 = cd{aw[0].x, au[0].x};
            l2 = cd{av[0].x, 0.0};
        } else if (k < H) {
            const cd a = aw[k], b = au[k];
            l1 = cd{a.x - b.y, a.y + b.x};
            l2 = av[k];
        } else if (k > H) {
            const cd a = aw[N - k], b = au[N - k];
            l1 = cd{a.x + b.y, b.x - a.y};
            l2 = conj(av[N - k]);
        }
        lds[pidx(k)] = l1;
        lds[LS + pidx(k)] = l2;
    }
    __syncthreads();
    fft_lines<N, LS, 2, T, true>(lds, twN, 1);
    for (int x = threadIdx.x; x < N; x += T) {
        const cd l1 = lds[pidx(x)], l2 = lds[LS + pidx(x)];
        const size_t o = (size_t)z * N + x;
        fields[o] = float(l1.x);
        fields[(size_t)N * N + o] = float(l1.y);
        fields[(size_t)2 * N * N + o] = float(l2.x);
    }
}

constexpr int kPackBlocks = 1024;

// small-scale fields hp [3][N][N] and per-block sums of the small-scale vorticity squared
__global__ __launch_bounds__(256) void k_closure_hp(const float *__restrict__ fields, int N, float *__restrict__ hp,
                                                    double *__restrict__ part)
{
    __shared__ double red[4];
    const size_t nn = (size_t)N * N;
    double s = 0.0;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < nn; p += (size_t)gridDim.x * 256) {
        const int y = (int)(p / N), x = (int)(p % N);
        const float w = closure_scaling::highpass(fields, N, y, x);
        hp[p] = w;
        hp[nn + p] = closure_scaling::highpass(fields + nn, N, y, x);
        hp[2 * nn + p] = closure_scaling::highpass(fields + 2 * nn, N, y, x);
        s += double(w) * double(w);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// scale[0] = 1 / w_s, scale[1] = 1 / (w_s D), scale[2] = w_s^2
__global__ __launch_bounds__(256) void k_closure_scale(const double *__restrict__ part, int blocks, int N,
                                                       double *__restrict__ scale)
{
    __shared__ double red[4];
    double s = 0.0;
    for (int b = threadIdx.x; b < blocks; b += 256) s += part[b];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double ws = sqrt((red[0] + red[1] + red[2] + red[3]) / (double(N) * N));
        const double D = 6.283185307179586476925 / N;
        if (ws > 1e-30) {
            scale[0] = 1.0 / ws;
            scale[1] = 1.0 / (ws * D);
            scale[2] = ws * ws;
        } else {
            scale[0] = scale[1] = scale[2] = 0.0;
        }
    }
}

__global__ void k_closure_pack(const float *__restrict__ hp, const double *__restrict__ scale, int N,
                               uint16_t *__restrict__ a0)
{
    const size_t nn = (size_t)N * N;
    const float s0 = float(scale[0]), s1 = float(scale[1]);
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < nn; p += (size_t)gridDim.x * blockDim.x) {
        uint4 v{0u, 0u, 0u, 0u};
        v.x = (uint32_t)f2bf(hp[p] * s0) | ((uint32_t)f2bf(hp[nn + p] * s1) << 16);
        v.y = (uint32_t)f2bf(hp[2 * nn + p] * s1);
        *reinterpret_cast<uint4 *>(a0 + p * 8) = v;
    }
}

template <int N, int T>
__global__ __launch_bounds__(T) void k_from_grid_x(const float *__restrict__ y, const double *__restrict__ scale,
                                                   const cd *__restrict__ twN, cd *__restrict__ bz)
{
    constexpr int H = N / 2, LS = padded_len(N);
    __shared__ cd lds[LS];
    const int q = blockIdx.x;
    const double s2 = scale[2];
    for (int x = threadIdx.x; x < N; x += T)
        lds[pidx(x)] = cd{s2 * y[(size_t)(2 * q) * N + x], s2 * y[(size_t)(2 * q + 1) * N + x]};
    __syncthreads();
    fft_lines<N, LS, 1, T, false>(lds, twN, 1);
    for (int kx = threadIdx.x; kx < H; kx += T) {
        const cd zk = lds[pidx(kx)], zm = conj(lds[pidx(kx == 0 ? 0 : N - kx)]);
        const cd sm = add(zk, zm), df = sub(zk, zm);
        bz[(size_t)(2 * q) * H + kx] = cd{0.5 * sm.x, 0.5 * sm.y};
        bz[(size_t)(2 * q + 1) * H + kx] = cd{0.5 * df.y, -0.5 * df.x};
    }
}

template <int N, int T>
__global__ __launch_bounds__(T) void k_from_grid_z(const cd *__restrict__ bz, const cd *__restrict__ twN,
                                                   cd *__restrict__ fext, ClosureShape sh)
{
    constexpr int H = N / 2, LS = padded_len(N);
    __shared__ cd lds[kNL * LS];
    const int kx0 = blockIdx.x * kNL;
    for (int e = threadIdx.x; e < kNL * N; e += T) {
        const int c = e % kNL, z = e / kNL;
        lds[c * LS + pidx(z)] = bz[(size_t)z * H + kx0 + c];
    }
    __syncthreads();
    fft_lines<N, LS, kNL, T, false>(lds, twN, 1);
    const double inv = 1.0 / (double(N) * double(N));
    for (int e = threadIdx.x; e < kNL * N; e += T) {
        const int c = e / N, iz = e - c * N;
        const int kx = kx0 + c;
        const double gz = (iz < H) ? double(iz) : double(iz - N);
        const double k = sqrt(double(kx) * kx + gz * gz);
        // band weight: 0 up to k_lo, 1 from k_hi, linear between; the mean (k = 0) and the
        // kz = -N/2 row always 0
        double wgt = (k <= sh.k_lo) ? 0.0 : (k >= sh.k_hi ? 1.0 : (k - sh.k_lo) / (sh.k_hi - sh.k_lo));
        if (iz == H || (kx == 0 && iz == 0)) wgt = 0.0;
        const cd v = lds[c * LS + pidx(iz)];
        cd f{v.x * inv * wgt, v.y * inv * wgt};
        if (sh.om) {
            // dissipative limiter: keep the part of f along -omega-hat, as an eddy damping
            // f = -nu omega-hat with 0 <= nu dt <= nu_dt_max
            const cd o = sh.om[(size_t)kx * N + iz];
            const double o2 = o.x * o.x + o.y * o.y;
            double nudt = o2 > 0.0 ? -(f.x * o.x + f.y * o.y) / o2 * sh.dt : 0.0;
            nudt = nudt < 0.0 ? 0.0 : (nudt > sh.nu_dt_max ? sh.nu_dt_max : nudt);
            const double nu = sh.dt > 0.0 ? nudt / sh.dt : 0.0;
            f = cd{-nu * o.x, -nu * o.y};
        }
        fext[(size_t)kx * N + iz] = f;
    }
}

template <int N>
hipError_t to_grid(const cd *om, const cd *twN, cd *az, float *fields, hipStream_t s)
{
    constexpr int T = N >= 256 ? 256 : 64;
    hipLaunchKernelGGL((k_to_grid_z<N, T>), dim3(N / 2 / kNL, 3), dim3(T), 0, s, om, twN, az);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_to_grid_x<N, T>), dim3(N), dim3(T), 0, s, az, twN, fields);
    return hipGetLastError();
}

template <int N>
hipError_t from_grid(const float *y, const double *scale, const cd *twN, cd *bz, cd *fext, const ClosureShape &sh,
                     hipStream_t s)
{
    constexpr int T = N >= 256 ? 256 : 64;
    hipLaunchKernelGGL((k_from_grid_x<N, T>), dim3(N / 2), dim3(T), 0, s, y, scale, twN, bz);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_from_grid_z<N, T>), dim3(N / 2 / kNL), dim3(T), 0, s, bz, twN, fext, sh);
    return hipGetLastError();
}

}  // namespace

int closure_grid_supported(int N)
{
    switch (N) {
        case 32: case 64: case 128: case 256: case 512: case 1024: return 1;
        default: return 0;
    }
}

hipError_t closure_launch_to_grid(const cd *om, int N, const cd *twN, cd *az, float *fields, hipStream_t s)
{
    switch (N) {
        case 32: return to_grid<32>(om, twN, az, fields, s);
        case 64: return to_grid<64>(om, twN, az, fields, s);
        case 128: return to_grid<128>(om, twN, az, fields, s);
        case 256: return to_grid<256>(om, twN, az, fields, s);
        case 512: return to_grid<512>(om, twN, az, fields, s);
        case 1024: return to_grid<1024>(om, twN, az, fields, s);
        default: return hipErrorInvalidValue;
    }
}

int closure_pack_blocks() { return kPackBlocks; }

hipError_t closure_launch_pack(const float *fields, int N, float *hp, double *part, double *scale, uint16_t *a0,
                               hipStream_t s)
{
    const size_t nn = (size_t)N * N;
    const int blocks = (int)((nn + 255) / 256 < (size_t)kPackBlocks ? (nn + 255) / 256 : (size_t)kPackBlocks);
    hipLaunchKernelGGL(k_closure_hp, dim3(blocks), dim3(256), 0, s, fields, N, hp, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_closure_scale, dim3(1), dim3(256), 0, s, part, blocks, N, scale);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_closure_pack, dim3(blocks), dim3(256), 0, s, hp, scale, N, a0);
    return hipGetLastError();
}

hipError_t closure_launch_from_grid(const float *y, const double *scale, int N, const cd *twN, cd *bz, cd *fext,
                                    const ClosureShape &sh, hipStream_t s)
{
    switch (N) {
        case 32: return from_grid<32>(y, scale, twN, bz, fext, sh, s);
        case 64: return from_grid<64>(y, scale, twN, bz, fext, sh, s);
        case 128: return from_grid<128>(y, scale, twN, bz, fext, sh, s);
        case 256: return from_grid<256>(y, scale, twN, bz, fext, sh, s);
        case 512: return from_grid<512>(y, scale, twN, bz, fext, sh, s);
        case 1024: return from_grid<1024>(y, scale, twN, bz, fext, sh, s);
        default: return hipErrorInvalidValue;
    }
}
