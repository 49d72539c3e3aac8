// dns_big.hip -- the time step for grids whose lines do not fit the fused LDS-resident
// kernels (N = 8192, 16384, 29400: 44100^2 padded), as seven passes through HBM.
//
// Every M-point transform is a four-step pair of LDS passes, M = L1 x L2:
//   pass A (L1-point FFTs over n1 of the points n1 L2 + n2, times W_M^{-/+ n2 k1})
//   pass B (L2-point FFTs over n2 of the pass-A results of one k1 -> point k1 + L1 k2)
// The state om, fn is [iz][kx] (rows of H = N/2 modes, the host layout), and every
// intermediate is row-major with kx (or x) contiguous, so the z-direction passes read
// and write runs of BZ adjacent kx and no transpose is needed anywhere.  One step
// (reference order mem_cuda.cu:1851-1857; advective form of FN):
//   1 zinv_a   om -> spectra of (u, w) and (d_x omega, d_z omega) on the padded z points,
//              z pass A                                                   Y [f][z1][kx]
//   2 zinv_b   z pass B -> the 4 fields on the physical rows z             Z [f][z][kx]
//   3 xinv_a   Hermitian row spectra u + i w and w_x + i w_z built on the fly from Z,
//              x pass A                                                    X [p][z][x1]
//   4 xinv_b   x pass B of rows 2q, 2q+1 (both pairs) -> physical values, product
//              P = u w_x + w w_z, the two rows packed as P_2q + i P_2q+1, and the first
//              forward step (L2-point FFT over k2 of each k1 line, x W_M^{k1 m2})  Yq [q][k1][m2]
//   5 xfwd_b   L1-point FFTs over k1 -> the packed rows' spectra, kept for kx < H and
//              M - kx                                                     Qh [q][kx]
//   6 zfwd_a   the two rows' spectra split apart (F_2q, F_2q+1) while loading, z pass A
//                                                                          W [z1][kx]
//   7 zfwd_b   z pass B, FN = -X / M^2 on the retained kz (kz = -N/2 row 0), CN/AB2
//              update of om, fn (step3_cuda.cu:75-100)
// The physical x order after pass 4 is (k1, k2) instead of k1 + L1 k2; the forward
// transform of passes 4-5 is written for that order, so nothing is reordered.
//
// Diagnostics (fields, CFL) reuse passes 1-2 in FP64, build natural-order rows and run
// plain batched four-step line FFTs (k_line_a / k_line_b).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>

#include "dns_big.h"
#include "fft_lds.hpp"

using namespace dnsfft;

namespace {

// threads per workgroup: the diagnostics' line FFTs (kT) and the step's passes (kTZ: the
// z-direction passes 1, 2, 6; kTXA: pass 3; kTXB: pass 4; kTXF: pass 5; kTZ7: pass 7),
// chosen by timing at N = 29400 (pass 7 at 256 threads instead of 512: 7.9 -> 7.4 ms FP64,
// 5.9 -> 5.8 ms with FP32 intermediates, same box; pass 5 at 256 instead of 512: 5.45 ->
// 4.55 ms FP64, 2.88 -> 2.55 ms FP32, 10.8 -> 9.15 ms in the u_i u_j form, while passes 1-3
// and 6 at 256 were slower; z tiles of 128 bytes, bz_of, timed against 64 and 256 bytes:
// both slower in every z pass)
constexpr int kT = 256;
constexpr int kTZ = 512, kTXA = 512, kTXB = 256, kTXF = 256, kTZ7 = 256;
// pass 7 of the u_i u_j form (two fields, 64 KB of LDS per workgroup: two workgroups per CU)
// at 512 threads: 9.74 -> 9.55 ms FP64 at N = 29400 (1024 threads: 10.8 ms)
constexpr int kTZ7U = 512;
// pass 3 with FP32 intermediates at N = 29400 runs 1024 threads (13.8 -> 12.6 ms; at
// N = 8192 and 16384 it is slower, and the other FP32 passes at 256 or 1024 threads were
// slower or equal)
template <class TC, int M_>
constexpr int txa_of()
{
    return (!std::is_same<TC, cd>::value && M_ == 44100) ? 1024 : kTXA;
}
// radix plan of pass 4's L2-point lines: radices up to 16 (N = 8192: 128 = 16 x 8 instead
// of 8 x 8 x 2, pass 4 1.25 -> 1.23 ms; at N = 16384 and 29400 the plan is the same either way)
constexpr int kXBR = 16;
// passes whose line transforms take the opaque thread index (fft_lines OPQ), where that
// timed faster: passes 4 and 7 with FP32 intermediates at N = 29400 (pass 4 11.6 -> 11.25 ms,
// pass 7 7.3 -> 6.3-7.1 ms; results bitwise unchanged); FP64, the other passes and the
// smaller grids (pass 4 at N = 16384: 2.97 -> 3.12 ms) are not faster with it
template <class TC, int M_>
constexpr bool big_opq(int pass)
{
    return !std::is_same<TC, cd>::value && M_ == 44100 && (pass == 4 || pass == 7);
}

__device__ __forceinline__ double gam(int iz, int N) { return (iz < (N >> 1)) ? double(iz) : double(iz - N); }

// Line lengths M = 3N/2 built for the multi-pass path: the sizes the parity tests force
// through it (N = 64 .. 1024), the reference's large grids (N = 8192, 9216, 16384, 20480:
// dns_cuda.cu:309-310, color_cuda.cu:493) and the 7-smooth family around them (N = 3072,
// 6144, 10240, 12288, 24576, 29400).  Everything below -- the split, the pass tiles and
// the size dispatch -- is derived from this one list.
constexpr int kBigM[] = {96,   192,   384,   768,   1536,  4608,  6144,  9216,
                         12288, 13824, 15360, 18432, 24576, 30720, 36864, 44100};
constexpr int kNumBigM = sizeof(kBigM) / sizeof(kBigM[0]);

constexpr bool smooth7_m(int m)
{
    for (int p : {2, 3, 5, 7})
        while (m % p == 0) m /= p;
    return m == 1;
}
// split M = L1 x L2 with L1 the largest divisor of M not above sqrt(M) (the squarest
// split: both line FFTs stay small enough for LDS), and the pass tiles C4 | L1 (k1 lines
// per pass-4 workgroup) and C5 | L2 (columns per pass-5 workgroup)
constexpr int split_l1(int m)
{
    int best = 1;
    for (int d = 1; d * d <= m; ++d)
        if (m % d == 0) best = d;
    return best;
}
constexpr int split_c4(int l1) { return l1 % 4 == 0 ? 4 : (l1 % 3 == 0 ? 3 : (l1 % 2 == 0 ? 2 : 1)); }
constexpr int split_c5(int l2)
{
    return l2 % 8 == 0 ? 8 : (l2 % 14 == 0 ? 14 : (l2 % 4 == 0 ? 4 : (l2 % 6 == 0 ? 6 : 2)));
}
template <int M_>
struct Split {
    static constexpr int L1 = split_l1(M_), L2 = M_ / L1, C4 = split_c4(L1), C5 = split_c5(L2);
    static_assert(smooth7_m(M_) && L1 * L2 == M_ && L2 % 2 == 0 && L1 % C4 == 0 && L2 % C5 == 0, "split");
    static_assert(L1 <= 256 && L2 <= 256, "line FFTs of at most 256 points (LDS)");
};
// M = 44100 = 210 x 210 (pass 4: 3 lines, pass 5: 14 columns), as timed
static_assert(Split<44100>::L1 == 210 && Split<44100>::C4 == 3 && Split<44100>::C5 == 14, "44100 split");

// f(std::integral_constant<int, M>) for the built M equal to m (hipErrorInvalidValue if none)
template <class F, int I = 0>
hipError_t with_big_m(int m, F &&f)
{
    if constexpr (I < kNumBigM) {
        if (m == kBigM[I]) return f(std::integral_constant<int, kBigM[I]>{});
        return with_big_m<F, I + 1>(m, static_cast<F &&>(f));
    } else {
        return hipErrorInvalidValue;
    }
}

// adjacent kx per z-direction workgroup: 128 bytes of elements
template <class TC>
constexpr int bz_of() { return sizeof(TC) == 16 ? 8 : 16; }

// Y (pass 1 -> 2) and W (pass 6 -> 7), which only the z-direction passes touch, are stored
// in column tiles of BZ modes, [kx / BZ][z][kx mod BZ] (each field M x hs elements as before):
// a z-pass workgroup (one tile, all rows) then stays within M x 128 bytes (5.6 MB at
// N = 29400, a few pages), where rows of stride hs (235 KB) spread its 210 rows over 49 MB
// and a third of its address translations missed the L1 TLB.
// The state om, fn is stored the same way on this path (dns_state_pos, dns_big.h): passes 1
// and 7 then read and write it within one tile's N x 128 bytes.
template <class TC>
__device__ __forceinline__ size_t ztile(int z, int kx, int M)
{
    constexpr int BZ = bz_of<TC>();
    return (size_t)(kx / BZ) * ((size_t)M * BZ) + (size_t)z * BZ + (kx % BZ);
}

// spectrum multiplier of field f on omega-hat (the fused column kernel's conventions:
// the multipliers listed with DNS_F_*)
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

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  Kernels
// Launch xcd_blocks(n) blocks for n logical blocks and take logical block xcd_order(b, n):
// workgroups go round-robin over the 8 XCDs (each with its own L2), so XCD x = b mod 8 runs
// the consecutive logical blocks [x per, (x + 1) per) -- neighbouring tiles, which share
// 128-byte lines at their edges, meet in one L2.  The grid is rounded up to whole sets of 8
// and the padding blocks (-1) return at once; without the padding a grid of n blocks with n
// mod 8 != 0 ran unmapped, with neighbours on different XCDs, and the passes that read lines
// shared by neighbouring tiles fetched them twice (x-forward B 1.26x, z-forward A 1.25x of
// their algorithmic bytes).
__host__ __device__ constexpr unsigned xcd_blocks(unsigned n) { return (n + 7) / 8 * 8; }
__device__ __forceinline__ int xcd_order(int b, int n)
{
    const int per = (n + 7) / 8, lb = (b % 8) * per + b / 8;
    return lb < n ? lb : -1;
}

// W_M^{e} from the table, conjugated for the inverse direction
template <class TC, bool INV>
__device__ __forceinline__ TC twid(const TC *__restrict__ tw, int e)
{
    const TC w = tw[opaque(e)];
    return INV ? conj(w) : w;
}

// Twiddles in LDS.  The passes take every twiddle from small LDS tables filled at kernel
// start from the W_M table, so no global load inside a transform waits behind the pass's
// own data loads (and the inter-step factors no longer miss in L1):
//   W_L^j of an L-point line FFT (L1: tw[j L2]; L2: tw[j L1]), L entries;
//   the inter-step factor W_M^e, e = n k < M, as W_M^{(e / L1) L1} W_M^{e % L1}: the
//   tables W_L2^a = tw[a L1] (a < L2, the L2 line table) and W_M^b = tw[b] (b < L1).
// A thread fetches its table entry into a register before issuing its data loads and
// writes it to LDS after them, so the table and data loads are in flight together.
template <class TC, int N, int T>
struct TabFetch {
    static_assert(N <= T, "one table entry per thread");
    TC v;
    __device__ __forceinline__ void fetch(const TC *__restrict__ tw, int stride)
    {
        v = threadIdx.x < N ? tw[threadIdx.x * stride] : TC{0, 0};
    }
    __device__ __forceinline__ void put(TC *dst) const
    {
        if (threadIdx.x < N) dst[threadIdx.x] = v;
    }
};

template <class TC, int L1, bool INV>
__device__ __forceinline__ TC twid2(const TC *wa, const TC *wb, int e)
{
    const TC w = mul(wa[e / L1], wb[e % L1]);
    return INV ? conj(w) : w;
}

// ---- pass 1: om -> two field spectra on the padded z points, z pass A ------------------
template <class TC, int M_, int T>
__global__ __launch_bounds__(T) void k_zinv_a(const cd *__restrict__ om, TC *__restrict__ Y, const TC *__restrict__ tw,
                                               BigFields fl, int H, size_t hs)
{
    using S = Split<M_>;
    constexpr int L1 = S::L1, L2 = S::L2, BZ = bz_of<TC>(), LS = padded_len(L1);
    __shared__ TC lds[2 * BZ * LS];
    __shared__ TC w1[L1], w2[L2], wb[L1];
    const int N = 2 * H;
    const int tiles = (H + BZ - 1) / BZ;
    // the field pairs of one (tile, n2) are neighbouring blocks: the second reads its om
    // values from the L2 (with the pair slowest, om was read from HBM once per pair)
    const int lb = xcd_order(blockIdx.x, L2 * tiles * fl.pairs);
    if (lb < 0) return;
    const int pr = lb % fl.pairs, rest = lb / fl.pairs;
    const int kx0 = (rest % tiles) * BZ, n2 = (rest / tiles) % L2;
    const int fa = fl.f[2 * pr], fb = fl.f[2 * pr + 1];
    // all of a thread's loads first (independent, in flight together), then the LDS writes
    constexpr int CNT = BZ * L1, IT = (CNT + T - 1) / T;
    TabFetch<TC, L1, T> tf1;
    TabFetch<TC, L2, T> tf2;
    TabFetch<TC, L1, T> tfb;
    tf1.fetch(tw, L2);
    tf2.fetch(tw, L1);
    tfb.fetch(tw, 1);
    cd ov[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * T;
        const int c = e % BZ, n1 = e / BZ;
        const int kx = kx0 + c, zp = n1 * L2 + n2;
        const int iz = zp < H ? zp : (zp >= 2 * H ? zp - H : -1);
        ov[i] = (e < CNT && kx < H && iz >= 0) ? om[dns_state_pos(iz, kx, N)] : cd{0.0, 0.0};
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * T;
        if (e < CNT) {
            const int c = e % BZ, n1 = e / BZ;
            const int kx = kx0 + c, zp = n1 * L2 + n2;
            const int iz = zp < H ? zp : (zp >= 2 * H ? zp - H : -1);
            TC va{0, 0}, vb{0, 0};
            if (kx < H && iz >= 0) {
                va = from_d<TC>(spec_of(fa, ov[i], kx, iz, N));
                vb = from_d<TC>(spec_of(fb, ov[i], kx, iz, N));
            }
            lds[c * LS + pidx(n1)] = va;
            lds[(BZ + c) * LS + pidx(n1)] = vb;
        }
    }
    tf1.put(w1);
    tf2.put(w2);
    tfb.put(wb);
    __syncthreads();
    fft_lines<L1, LS, 2 * BZ, T, true>(lds, w1, 1);
    TC *ya = Y + (size_t)(2 * pr) * M_ * hs, *yb = ya + (size_t)M_ * hs;
    for (int e = threadIdx.x; e < BZ * L1; e += T) {
        const int c = e % BZ, k1 = e / BZ;
        const int kx = kx0 + c;
        if (kx < H) {
            const TC w = twid2<TC, L1, true>(w2, wb, n2 * k1);
            const size_t o = ztile<TC>(k1 * L2 + n2, kx, M_);
            ya[o] = mul(lds[c * LS + pidx(k1)], w);
            yb[o] = mul(lds[(BZ + c) * LS + pidx(k1)], w);
        }
    }
}

// ---- pass 2: z pass B -> fields on the physical rows ------------------------------------
template <class TC, int M_, int T>
__global__ __launch_bounds__(T) void k_zinv_b(const TC *__restrict__ Y, TC *__restrict__ Z, const TC *__restrict__ tw,
                                               int H, size_t hs, int pairs)
{
    using S = Split<M_>;
    constexpr int L1 = S::L1, L2 = S::L2, BZ = bz_of<TC>(), LS = padded_len(L2);
    __shared__ TC lds[2 * BZ * LS];
    __shared__ TC w2[L2];
    const int tiles = (H + BZ - 1) / BZ;
    const int lb = xcd_order(blockIdx.x, L1 * tiles * pairs);
    if (lb < 0) return;
    const int kx0 = (lb % tiles) * BZ, k1 = (lb / tiles) % L1, pr = lb / (tiles * L1);
    const TC *ya = Y + (size_t)(2 * pr) * M_ * hs, *yb = ya + (size_t)M_ * hs;
    constexpr int CNT = BZ * L2, IT = (CNT + T - 1) / T;
    TabFetch<TC, L2, T> tf2;
    tf2.fetch(tw, L1);
    TC va[IT], vb[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * T;
        const int c = e % BZ, n2 = e / BZ;
        const int kx = kx0 + c;
        const size_t o = ztile<TC>(k1 * L2 + n2, kx, M_);
        const bool ok = e < CNT && kx < H;
        va[i] = ok ? ya[o] : TC{0, 0};
        vb[i] = ok ? yb[o] : TC{0, 0};
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * T;
        if (e < CNT) {
            const int c = e % BZ, n2 = e / BZ;
            lds[c * LS + pidx(n2)] = va[i];
            lds[(BZ + c) * LS + pidx(n2)] = vb[i];
        }
    }
    tf2.put(w2);
    __syncthreads();
    fft_lines<L2, LS, 2 * BZ, T, true>(lds, w2, 1);
    TC *za = Z + (size_t)(2 * pr) * M_ * hs, *zb = za + (size_t)M_ * hs;
    for (int e = threadIdx.x; e < BZ * L2; e += T) {
        const int c = e % BZ, k2 = e / BZ;
        const int kx = kx0 + c;
        if (kx < H) {
            const size_t o = (size_t)(k1 + L1 * k2) * hs + kx;
            za[o] = lds[c * LS + pidx(k2)];
            zb[o] = lds[(BZ + c) * LS + pidx(k2)];
        }
    }
}

// Hermitian row spectrum a + i b at kx from two fields' row (kx < H retained; bin 0 keeps
// the real parts -- the C2R convention; M - kx mirrors; the band between is zero)
template <class TC>
__device__ __forceinline__ TC row_spec(const TC *__restrict__ a, const TC *__restrict__ b, int kx, int H, int M)
{
    if (kx == 0) return TC{a[0].x, b[0].x};
    if (kx < H) {
        const TC av = a[kx], bv = b[kx];
        return TC{av.x - bv.y, av.y + bv.x};
    }
    if (kx > M - H) {
        const TC av = a[M - kx], bv = b[M - kx];
        return TC{av.x + bv.y, bv.x - av.y};
    }
    return TC{0, 0};
}

// ---- pass 3: row spectra built from Z, x pass A ------------------------------------------
// A workgroup owns a mirror-closed set of columns n2 of one row and one plane p (p = 0:
// u + i w from fields 0, 1; p = 1: w_x + i w_z from fields 2, 3): the pairs
// {j, L2 - j} for CP consecutive j in [1, L2/2), or (the last tile) the self-mirrored
// columns 0 and L2/2.  Positions kx and M - kx then both lie in its lines, so each
// retained value of the row is read once and fills its direct slot (a + i b) and its
// mirror slot (conj a + i conj b); the band H <= kx <= M - H is zero.
template <class TC, int M_, int T>
__global__ __launch_bounds__(T) void k_xinv_a(const TC *__restrict__ Z, TC *__restrict__ X, const TC *__restrict__ tw,
                                               int H, size_t hs, size_t ms, int planes)
{
    using S = Split<M_>;
    constexpr int L1 = S::L1, L2 = S::L2, CP = bz_of<TC>(), LS = padded_len(L1);
    constexpr int NPT = (L2 / 2 - 1 + CP - 1) / CP;  // pair tiles; tile NPT holds columns 0, L2/2
    __shared__ TC lds[2 * CP * LS];
    __shared__ TC w1[L1], w2[L2], wb[L1];
    const int lb = xcd_order(blockIdx.x, (NPT + 1) * M_ * planes);
    if (lb < 0) return;
    const int t = lb % (NPT + 1), z = (lb / (NPT + 1)) % M_, p = lb / ((NPT + 1) * M_);
    const TC *a = Z + ((size_t)(2 * p) * M_ + z) * hs;
    const TC *b = a + (size_t)M_ * hs;
    constexpr int CNT = CP * L1, IT = (CNT + T - 1) / T;
    TabFetch<TC, L1, T> tf1;
    TabFetch<TC, L2, T> tf2;
    TabFetch<TC, L1, T> tfb;
    tf1.fetch(tw, L2);
    tf2.fetch(tw, L1);
    tfb.fetch(tw, 1);
    TC va[IT], vb[IT];
    if (t < NPT) {
        // slot pair (line i, n1) = position kp = n1 L2 + d, d = 1 + t CP + i, and (line CP + i,
        // L1 - 1 - n1) = position M - kp; the smaller of the two is the retained one
#pragma unroll
        for (int k = 0; k < IT; ++k) {
            const int e = threadIdx.x + k * T;
            const int i = e % CP, n1 = e / CP;
            const int d = 1 + t * CP + i;
            const int kp = n1 * L2 + d, lo = kp < M_ - kp ? kp : M_ - kp;
            const bool ok = e < CNT && d < L2 / 2 && lo < H;
            va[k] = ok ? a[lo] : TC{0, 0};
            vb[k] = ok ? b[lo] : TC{0, 0};
        }
#pragma unroll
        for (int k = 0; k < IT; ++k) {
            const int e = threadIdx.x + k * T;
            if (e < CNT) {
                const int i = e % CP, n1 = e / CP;
                const int kp = n1 * L2 + 1 + t * CP + i;
                const TC av = va[k], bv = vb[k];
                const TC dir{av.x - bv.y, av.y + bv.x}, mir{av.x + bv.y, bv.x - av.y};
                const bool first = kp < M_ - kp;  // this slot holds the retained position
                // component-wise selects (a select of the 16-byte structs goes through scratch)
                lds[i * LS + pidx(n1)] = TC{first ? dir.x : mir.x, first ? dir.y : mir.y};
                lds[(CP + i) * LS + pidx(L1 - 1 - n1)] = TC{first ? mir.x : dir.x, first ? mir.y : dir.y};
            }
        }
    } else {
        // columns 0 and L2/2 (lines 0, 1), each its own mirror; other lines unused
        constexpr int CNT2 = 2 * L1, IT2 = (CNT2 + T - 1) / T;
#pragma unroll
        for (int k = 0; k < IT2; ++k) {
            const int e = threadIdx.x + k * T;
            const int col = (e / L1) * (L2 / 2), n1 = e % L1;
            const int kp = n1 * L2 + col;
            const int src = kp < H ? kp : (kp > M_ - H ? M_ - kp : -1);
            TC v{0, 0};
            if (e < CNT2 && src >= 0) {
                const TC av = a[src], bv = b[src];
                if (kp == 0)
                    v = TC{av.x, bv.x};
                else if (kp < H)
                    v = TC{av.x - bv.y, av.y + bv.x};
                else
                    v = TC{av.x + bv.y, bv.x - av.y};
            }
            if (e < CNT2) lds[(e / L1) * LS + pidx(n1)] = v;
        }
        for (int e = threadIdx.x; e < (2 * CP - 2) * L1; e += T) lds[(2 + e / L1) * LS + pidx(e % L1)] = TC{0, 0};
    }
    tf1.put(w1);
    tf2.put(w2);
    tfb.put(wb);
    __syncthreads();
    fft_lines<L1, LS, 2 * CP, T, true>(lds, w1, 1);
    const int nl = t < NPT ? 2 * CP : 2;
    for (int e = threadIdx.x; e < nl * L1; e += T) {
        const int l = e % nl, k1 = e / nl;
        int n2;
        if (t < NPT) {
            const int d = 1 + t * CP + (l % CP);
            if (d >= L2 / 2) continue;
            n2 = l < CP ? d : L2 - d;
        } else {
            n2 = l * (L2 / 2);
        }
        X[((size_t)p * M_ + z) * ms + (size_t)k1 * L2 + n2] = mul(lds[l * LS + pidx(k1)], twid2<TC, L1, true>(w2, wb, n2 * k1));
    }
}

// ---- pass 4: x pass B of a row pair, product, first forward step -------------------------
// row pairs per pass-4 workgroup: two in the u_i u_j form and the CFL form (the tables and
// barriers of a workgroup then serve twice the lines), one in the advective form (whose four
// groups per pair already fill the LDS of three workgroups per CU)
template <int M_, int MODE>
constexpr int xb_pairs() { return MODE != 0 && M_ % 4 == 0 ? 2 : 1; }
// MODE 0: product -> Yq;  MODE 1: CFL maxima of u, w (plane 0 only) -> cflpart
template <class TC, int M_, int MODE, int T>
__global__ __launch_bounds__(T) void k_xinv_b(const TC *__restrict__ X, TC *__restrict__ Yq, const TC *__restrict__ tw,
                                               size_t ms, double *__restrict__ cflpart, double DX, double DZ)
{
    using S = Split<M_>;
    constexpr int L1 = S::L1, L2 = S::L2, C = S::C4, LS = padded_len(L2);
    // lines per k1: group g = 2 p + r (plane p, row 2q + r); MODES 1 and 2 (CFL, u_i u_j form)
    // read plane 0 (u + i w) of RP row pairs (rows 2 RP q + g); MODE 2 forward-transforms two
    // products per pair
    constexpr int RP = xb_pairs<M_, MODE>();
    constexpr int G = MODE == 0 ? 4 : 2 * RP;
    __shared__ TC lds[G * C * LS];
    __shared__ TC w2[L2], wb[L1];
    __shared__ double red[T / 64];
    using RT = decltype(TC::x);
    const int q = blockIdx.x, k10 = blockIdx.y * C;
    // TPN: one thread per line position n2 (L2 <= T < 2 L2, e.g. 210 points at 256 threads),
    // every line's element n2 in that thread: the line, row and plane of each access are
    // compile-time constants, where the flat element loop pays two divisions per element
    constexpr bool TPN = L2 <= T && 2 * L2 > T;
    constexpr int CNT = G * C * L2, IT = TPN ? G * C : (CNT + T - 1) / T;
    TabFetch<TC, L2, T> tf2;
    TabFetch<TC, L1, T> tfb;
    tf2.fetch(tw, L1);
    tfb.fetch(tw, 1);
    TC xv[IT];
    const auto row_of = [&](int g) { return MODE == 0 ? (size_t)(g >> 1) * M_ + 2 * q + (g & 1) : (size_t)2 * RP * q + g; };
    if constexpr (TPN) {
        const int n2 = threadIdx.x;
        if (n2 < L2) {
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int c = 0; c < C; ++c) xv[g * C + c] = X[row_of(g) * ms + (size_t)(k10 + c) * L2 + n2];
#pragma unroll
            for (int l = 0; l < G * C; ++l) lds[l * LS + pidx(n2)] = xv[l];
        }
    } else {
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int e = threadIdx.x + i * T;
            const int n2 = e % L2, rest = e / L2, c = rest % C, g = rest / C;
            xv[i] = e < CNT ? X[row_of(g) * ms + (size_t)(k10 + c) * L2 + n2] : TC{0, 0};
        }
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int e = threadIdx.x + i * T;
            if (e < CNT) {
                const int n2 = e % L2, rest = e / L2, c = rest % C, g = rest / C;
                lds[(g * C + c) * LS + pidx(n2)] = xv[i];
            }
        }
    }
    tf2.put(w2);
    tfb.put(wb);
    __syncthreads();
    fft_lines<L2, LS, G * C, T, true, kXBR, big_opq<TC, M_>(4)>(lds, w2, 1);
    if (MODE == 1) {
        double m = 0.0;
        for (int e = threadIdx.x; e < G * C * L2; e += T) {
            const TC v = lds[(e / L2) * LS + pidx(e % L2)];
            const double cfl = fabs(double(v.x)) / DX + fabs(double(v.y)) / DZ;
            m = cfl > m ? cfl : m;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(m, off, 64);
            m = o > m ? o : m;
        }
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = red[0];
            for (int w = 1; w < T / 64; ++w) t = red[w] > t ? red[w] : t;
            cflpart[(size_t)q * gridDim.y + blockIdx.y] = t;
        }
        return;
    }
    if constexpr (MODE == 2) {
        // D = u^2 - w^2 and C = u w of rows 2j (real parts) and 2j + 1 (imaginary parts) of
        // pair j = RP q + pp: lines 2 pp C + c hold D, lines (2 pp + 1) C + c hold C; all
        // x-forwarded, written to Yq planes 0 (D) and 1 (C)
        const auto prod = [&](int l0, int k2) {
            const int l1 = l0 + C;
            const TC a0 = lds[l0 * LS + pidx(k2)], a1 = lds[l1 * LS + pidx(k2)];
            lds[l0 * LS + pidx(k2)] = TC{RT(a0.x * a0.x - a0.y * a0.y), RT(a1.x * a1.x - a1.y * a1.y)};
            lds[l1 * LS + pidx(k2)] = TC{RT(a0.x * a0.y), RT(a1.x * a1.y)};
        };
        if constexpr (TPN) {
            if (threadIdx.x < L2)
#pragma unroll
                for (int pp = 0; pp < RP; ++pp)
#pragma unroll
                    for (int c = 0; c < C; ++c) prod(2 * pp * C + c, threadIdx.x);
        } else {
            for (int e = threadIdx.x; e < RP * C * L2; e += T) {
                const int pc = e / L2, k2 = e % L2, pp = pc / C, c = pc - pp * C;
                prod(2 * pp * C + c, k2);
            }
        }
        __syncthreads();
        fft_lines<L2, LS, 2 * RP * C, T, false, kXBR, big_opq<TC, M_>(4)>(lds, w2, 1);
        // line g C + c: pair pp = g / 2, plane g mod 2 (D, C), k1 = k10 + c
        const auto out = [&](int g, int c, int m2, TC w) {
            Yq[((size_t)(g & 1) * (M_ / 2) + RP * q + (g >> 1)) * ms + (size_t)(k10 + c) * L2 + m2] =
                mul(lds[(g * C + c) * LS + pidx(m2)], w);
        };
        if constexpr (TPN) {
            const int m2 = threadIdx.x;
            if (m2 < L2)
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const TC w = twid2<TC, L1, false>(w2, wb, (k10 + c) * m2);  // one twiddle for the 2 RP lines
#pragma unroll
                    for (int g = 0; g < 2 * RP; ++g) out(g, c, m2, w);
                }
        } else {
            for (int e = threadIdx.x; e < 2 * RP * C * L2; e += T) {
                const int gc = e / L2, m2 = e % L2, g = gc / C, c = gc - g * C;
                out(g, c, m2, twid2<TC, L1, false>(w2, wb, (k10 + c) * m2));
            }
        }
        return;
    }
    // P = u w_x + w w_z on rows 2q (real part) and 2q + 1 (imaginary part): group 0/1 hold
    // u + i w of the two rows, group 2/3 w_x + i w_z
    const auto prod = [&](int c, int k2) {
        const TC a0 = lds[(0 * C + c) * LS + pidx(k2)], a1 = lds[(1 * C + c) * LS + pidx(k2)];
        const TC b0 = lds[(2 * C + c) * LS + pidx(k2)], b1 = lds[(3 * C + c) * LS + pidx(k2)];
        lds[c * LS + pidx(k2)] = TC{RT(a0.x * b0.x + a0.y * b0.y), RT(a1.x * b1.x + a1.y * b1.y)};
    };
    const auto out = [&](int c, int m2) {
        const int k1 = k10 + c;
        Yq[(size_t)q * ms + (size_t)k1 * L2 + m2] = mul(lds[c * LS + pidx(m2)], twid2<TC, L1, false>(w2, wb, k1 * m2));
    };
    if constexpr (TPN) {
        if (threadIdx.x < L2)
#pragma unroll
            for (int c = 0; c < C; ++c) prod(c, threadIdx.x);
    } else {
        for (int e = threadIdx.x; e < C * L2; e += T) prod(e / L2, e % L2);
    }
    __syncthreads();
    fft_lines<L2, LS, C, T, false, kXBR, big_opq<TC, M_>(4)>(lds, w2, 1);
    if constexpr (TPN) {
        if (threadIdx.x < L2)
#pragma unroll
            for (int c = 0; c < C; ++c) out(c, threadIdx.x);
    } else {
        for (int e = threadIdx.x; e < C * L2; e += T) out(e / L2, e % L2);
    }
}

// ---- pass 5: L1-point FFTs over k1 -> packed row-pair spectra (kept modes only) ----------
template <class TC, int M_, int T>
__global__ __launch_bounds__(T) void k_xfwd_b(const TC *__restrict__ Yq, TC *__restrict__ Qh, const TC *__restrict__ tw,
                                               int H, size_t ms, int rows)
{
    using S = Split<M_>;
    constexpr int L1 = S::L1, L2 = S::L2, C = S::C5, LS = padded_len(L1);
    __shared__ TC lds[C * LS];
    const int lb = xcd_order(blockIdx.x, rows * (L2 / C));  // rows: M / 2, or M for two planes
    if (lb < 0) return;
    const int m20 = (lb % (L2 / C)) * C, q = lb / (L2 / C);
    constexpr int CNT = C * L1, IT = (CNT + T - 1) / T;
    TC yv[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * T;
        const int c = e % C, k1 = e / C;
        yv[i] = e < CNT ? Yq[(size_t)q * ms + (size_t)k1 * L2 + m20 + c] : TC{0, 0};
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * T;
        if (e < CNT) lds[(e % C) * LS + pidx(e / C)] = yv[i];
    }
    __syncthreads();
    fft_lines<L1, LS, C, T, false>(lds, tw, L2);  // global table: L1-cached stage twiddles
    for (int e = threadIdx.x; e < C * L1; e += T) {
        const int c = e % C, m1 = e / C;
        const int kx = m20 + c + L2 * m1;
        if (kx < H || kx > M_ - H) Qh[(size_t)q * ms + kx] = lds[c * LS + pidx(m1)];
    }
}

// ---- pass 6: split the row pairs, z pass A (forward) -------------------------------------
template <class TC, int M_, int T>
__global__ __launch_bounds__(T) void k_zfwd_a(const TC *__restrict__ Qh, TC *__restrict__ W, const TC *__restrict__ tw,
                                               int H, size_t hs, size_t ms)
{
    using S = Split<M_>;
    constexpr int L1 = S::L1, L2 = S::L2, BZ = bz_of<TC>(), LS = padded_len(L1);
    __shared__ TC lds[2 * BZ * LS];
    __shared__ TC w1[L1], w2[L2], wb[L1];
    using RT = decltype(TC::x);
    const int tiles = (H + BZ - 1) / BZ;
    const int lb = xcd_order(blockIdx.x, (L2 / 2) * tiles);
    if (lb < 0) return;
    const int kx0 = (lb % tiles) * BZ, j = lb / tiles;  // n2 = 2 j + r: rows z of parity r
    const RT half = RT(0.5);
    constexpr int CNT = BZ * L1, IT = (CNT + T - 1) / T;
    TabFetch<TC, L1, T> tf1;
    TabFetch<TC, L2, T> tf2;
    TabFetch<TC, L1, T> tfb;
    tf1.fetch(tw, L2);
    tf2.fetch(tw, L1);
    tfb.fetch(tw, 1);
    TC zkv[IT], zmv[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * T;
        const int c = e % BZ, n1 = e / BZ;
        const int kx = kx0 + c;
        const size_t q = (size_t)(n1 * L2 + 2 * j) / 2;
        const bool ok = e < CNT && kx < H;
        zkv[i] = ok ? Qh[q * ms + kx] : TC{0, 0};
        zmv[i] = ok ? Qh[q * ms + (kx == 0 ? 0 : M_ - kx)] : TC{0, 0};
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * T;
        if (e < CNT) {
            const int c = e % BZ, n1 = e / BZ;
            const TC zk = zkv[i], zm = conj(zmv[i]);
            const TC sm = add(zk, zm), df = sub(zk, zm);
            lds[c * LS + pidx(n1)] = TC{half * sm.x, half * sm.y};
            lds[(BZ + c) * LS + pidx(n1)] = TC{half * df.y, -half * df.x};
        }
    }
    tf1.put(w1);
    tf2.put(w2);
    tfb.put(wb);
    __syncthreads();
    fft_lines<L1, LS, 2 * BZ, T, false>(lds, w1, 1);
    for (int e = threadIdx.x; e < 2 * BZ * L1; e += T) {
        const int c = e % BZ, rest = e / BZ, k1 = rest % L1, r = rest / L1;
        const int kx = kx0 + c, n2 = 2 * j + r;
        if (kx < H)
            W[ztile<TC>(k1 * L2 + n2, kx, M_)] = mul(lds[(r * BZ + c) * LS + pidx(k1)], twid2<TC, L1, false>(w2, wb, n2 * k1));
    }
}

// ---- pass 7: z pass B (forward), FN, CN/AB2 ----------------------------------------------
template <class TC, int M_, int T, bool UI = false>
__global__ __launch_bounds__(T) void k_zfwd_b(const TC *__restrict__ W, cd *__restrict__ om, cd *__restrict__ fn,
                                               const TC *__restrict__ tw, int H, size_t hs, BigCnArgs a)
{
    // UI: the u_i u_j form -- W holds the z-pass-A outputs of D = u^2 - w^2 (field 0) and C = u w
    // (field 1, M hs further); FN = (alpha gamma D^ + (gamma^2 - alpha^2) C^) / M^2
    using S = Split<M_>;
    constexpr int L1 = S::L1, L2 = S::L2, BZ = bz_of<TC>(), LS = padded_len(L2);
    constexpr int NF = UI ? 2 : 1;
    __shared__ TC lds[NF * BZ * LS];
    __shared__ TC w2[L2];
    const int N = 2 * H;
    const int tiles = (H + BZ - 1) / BZ;
    // plain block order: its tiles share no lines (XCD-mapped: 7.47 -> 7.63 ms at N = 29400)
    const int lb = blockIdx.x;
    const int kx0 = (lb % tiles) * BZ, k1 = lb / tiles;
    constexpr int CNT = NF * BZ * L2, IT = (CNT + T - 1) / T;
    TabFetch<TC, L2, T> tf2;
    tf2.fetch(tw, L1);
    TC wv[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * T;
        const int c = e % BZ, rest = e / BZ, n2 = rest % L2, f = rest / L2;
        const int kx = kx0 + c;
        wv[i] = (e < CNT && kx < H) ? W[(size_t)f * M_ * hs + ztile<TC>(k1 * L2 + n2, kx, M_)] : TC{0, 0};
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = threadIdx.x + i * T;
        if (e < CNT) {
            const int c = e % BZ, rest = e / BZ, n2 = rest % L2, f = rest / L2;
            lds[(f * BZ + c) * LS + pidx(n2)] = wv[i];
        }
    }
    tf2.put(w2);
    __syncthreads();
    fft_lines<L2, LS, NF * BZ, T, false, 8, big_opq<TC, M_>(7)>(lds, w2, 1);
    for (int e = threadIdx.x; e < BZ * L2; e += T) {
        const int c = e % BZ, k2 = e / BZ;
        const int kx = kx0 + c, kz = k1 + L1 * k2;
        const int iz = kz < H ? kz : (kz >= 2 * H ? kz - H : -1);
        if (kx >= H || iz < 0) continue;
        const double ax = double(kx), gz = gam(iz, N);
        const cd X = to_d(lds[c * LS + pidx(k2)]);
        cd FN{-X.x * a.divxz, -X.y * a.divxz};
        if constexpr (UI) {
            const cd Xc = to_d(lds[(BZ + c) * LS + pidx(k2)]);
            const double ga = gz * ax * a.divxz, gm = (gz * gz - ax * ax) * a.divxz;
            FN = cd{ga * X.x + gm * Xc.x, ga * X.y + gm * Xc.y};
        }
        if (iz == H) FN = cd{0.0, 0.0};
        const double ARG = a.VT * (ax * ax + gz * gz);
        const double DEN = 1.0 + ARG, c1 = 1.0 - ARG;
        const size_t io = dns_state_pos(iz, kx, N);
        const cd o = om[io], f = fn[io];
        const cd num{c1 * o.x + a.c2 * FN.x + a.c3 * f.x, c1 * o.y + a.c2 * FN.y + a.c3 * f.y};
        om[io] = cd{num.x / DEN, num.y / DEN};
        fn[io] = FN;
    }
}

// ---- diagnostics: natural-order rows and plain batched line FFTs (FP64) -------------------
// rows R[z][x] of the Hermitian spectrum of field 0 of Z (+ i field 1 when two)
__global__ void k_rows(const cd *__restrict__ Z, cd *__restrict__ R, int H, int M, size_t hs, int two)
{
    const size_t total = (size_t)M * M;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int z = (int)(e / M), x = (int)(e % M);
        const cd *a = Z + (size_t)z * hs;
        if (two) {
            R[e] = row_spec(a, a + (size_t)M * hs, x, H, M);
        } else {
            cd v{0.0, 0.0};
            if (x == 0)
                v = cd{a[0].x, 0.0};
            else if (x < H)
                v = a[x];
            else if (x > M - H)
                v = conj(a[M - x]);
            R[e] = v;
        }
    }
}

template <int L1, int L2, int C, bool INV>
__global__ __launch_bounds__(kT) void k_line_a(const cd *__restrict__ x, cd *__restrict__ y, const cd *__restrict__ tw)
{
    constexpr int L = L1 * L2, LS = padded_len(L1);
    __shared__ cd lds[C * LS];
    const size_t line = blockIdx.y;
    const int n20 = blockIdx.x * C;
    const cd *src = x + line * L;
    for (int e = threadIdx.x; e < C * L1; e += kT) {
        const int c = e % C, n1 = e / C;
        lds[c * LS + pidx(n1)] = src[(size_t)n1 * L2 + n20 + c];
    }
    __syncthreads();
    fft_lines<L1, LS, C, kT, INV>(lds, tw, L2);
    cd *dst = y + line * L;
    for (int e = threadIdx.x; e < C * L1; e += kT) {
        const int c = e % C, k1 = e / C;
        const int n2 = n20 + c;
        dst[(size_t)k1 * L2 + n2] = mul(lds[c * LS + pidx(k1)], twid<cd, INV>(tw, n2 * k1));
    }
}

template <int L1, int L2, int C, bool INV>
__global__ __launch_bounds__(kT) void k_line_b(const cd *__restrict__ y, cd *__restrict__ out, const cd *__restrict__ tw)
{
    constexpr int L = L1 * L2, LS = padded_len(L2);
    __shared__ cd lds[C * LS];
    const size_t line = blockIdx.y;
    const int k10 = blockIdx.x * C;
    const cd *src = y + line * L;
    for (int e = threadIdx.x; e < C * L2; e += kT) {
        const int c = e / L2, n2 = e % L2;
        lds[c * LS + pidx(n2)] = src[(size_t)(k10 + c) * L2 + n2];
    }
    __syncthreads();
    fft_lines<L2, LS, C, kT, INV>(lds, tw, L1);
    cd *dst = out + line * L;
    for (int e = threadIdx.x; e < C * L2; e += kT) {
        const int c = e % C, k2 = e / C;
        dst[(size_t)k2 * L1 + k10 + c] = lds[c * LS + pidx(k2)];
    }
}

__global__ void k_rows_phys(const cd *__restrict__ R, size_t n, double *__restrict__ out0, int out_mode)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const cd v = R[e];
        out0[e] = out_mode == 1 ? sqrt(v.x * v.x + v.y * v.y) : v.x;
    }
}

// E, Z partial sums per column tile of the state (the energy convention of k_energy): block
// b sums the modes kx = 8 b .. 8 b + 7 over all iz (one tile's contiguous N x 128 bytes)
__global__ __launch_bounds__(256) void k_energy_tiles(const cd *__restrict__ om, int N, double *__restrict__ pe,
                                                      double *__restrict__ pz)
{
    __shared__ double re[4], rz[4];
    const int H = N / 2, b = blockIdx.x;
    const cd *t = om + dns_state_pos(0, 8 * b, N);
    double e = 0.0, zs = 0.0;
    for (int i = threadIdx.x; i < 8 * N; i += 256) {
        const int iz = i >> 3, kx = 8 * b + (i & 7);
        if (kx >= H) continue;
        const cd o = t[i];
        const double gz = gam(iz, N);
        const double a2 = o.x * o.x + o.y * o.y, m = kx == 0 ? 1.0 : 2.0;
        const double k2 = double(kx) * kx + gz * gz;
        zs += m * a2;
        if (k2 > 0.0) e += m * a2 / k2;
    }
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off, 64);
        zs += __shfl_xor(zs, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        re[threadIdx.x >> 6] = e;
        rz[threadIdx.x >> 6] = zs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        pe[b] = re[0] + re[1] + re[2] + re[3];
        pz[b] = rz[0] + rz[1] + rz[2] + rz[3];
    }
}

// the state between the host layout [iz][kx] (N x H) and the tiles: to_tiles 1 reads rows
// from src and writes tiles to dst, 0 the other way
__global__ void k_state_tile(const cd *__restrict__ src, cd *__restrict__ dst, int N, int to_tiles)
{
    const int H = N / 2;
    const size_t total = (size_t)N * H;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(e / H), kx = (int)(e % H);
        const size_t t = dns_state_pos(iz, kx, N);
        if (to_tiles)
            dst[t] = src[e];
        else
            dst[e] = src[t];
    }
}

int grid_for(size_t n) { return (int)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536); }

template <class TC, int M_>
hipError_t launch_m(const BigReq &r)
{
    using S = Split<M_>;
    constexpr int BZ = bz_of<TC>();
    const int H = r.H;
    const unsigned tiles = (unsigned)((H + BZ - 1) / BZ);
    const TC *tw = static_cast<const TC *>(r.tw);
    hipStream_t s = r.stream;
    switch (r.what) {
        case BIG_ZINV_A:
            hipLaunchKernelGGL((k_zinv_a<TC, M_, kTZ>), dim3(xcd_blocks(S::L2 * tiles * (unsigned)r.fields.pairs)), dim3(kTZ), 0, s, r.om,
                               static_cast<TC *>(r.b), tw, r.fields, H, r.hs);
            break;
        case BIG_ZINV_B:
            hipLaunchKernelGGL((k_zinv_b<TC, M_, kTZ>), dim3(xcd_blocks(S::L1 * tiles * (unsigned)r.fields.pairs)), dim3(kTZ), 0, s,
                               static_cast<const TC *>(r.a), static_cast<TC *>(r.b), tw, H, r.hs, r.fields.pairs);
            break;
        case BIG_XINV_A: {
            constexpr int NPT = (S::L2 / 2 - 1 + BZ - 1) / BZ;
            hipLaunchKernelGGL((k_xinv_a<TC, M_, txa_of<TC, M_>()>), dim3(xcd_blocks((NPT + 1) * M_ * (unsigned)r.planes)), dim3(txa_of<TC, M_>()), 0, s,
                               static_cast<const TC *>(r.a), static_cast<TC *>(r.b), tw, H, r.hs, r.ms, r.planes);
            break;
        }
        case BIG_XINV_B:
            if (r.uiuj)
                hipLaunchKernelGGL((k_xinv_b<TC, M_, 2, kTXB>), dim3(M_ / (2 * xb_pairs<M_, 2>()), S::L1 / S::C4), dim3(kTXB), 0, s,
                                   static_cast<const TC *>(r.a), static_cast<TC *>(r.b), tw, r.ms, nullptr, 0.0, 0.0);
            else
                hipLaunchKernelGGL((k_xinv_b<TC, M_, 0, kTXB>), dim3(M_ / 2, S::L1 / S::C4), dim3(kTXB), 0, s,
                                   static_cast<const TC *>(r.a), static_cast<TC *>(r.b), tw, r.ms, nullptr, 0.0, 0.0);
            break;
        case BIG_XINV_CFL:
            hipLaunchKernelGGL((k_xinv_b<TC, M_, 1, kTXB>), dim3(M_ / (2 * xb_pairs<M_, 1>()), S::L1 / S::C4), dim3(kTXB), 0, s,
                               static_cast<const TC *>(r.a), nullptr, tw, r.ms, r.cflrow, r.DX, r.DZ);
            break;
        case BIG_XFWD_B:
        {
            // the u_i u_j form's two planes (D, C) are M / 2 rows each, back to back
            const int rows = r.uiuj ? M_ : M_ / 2;
            hipLaunchKernelGGL((k_xfwd_b<TC, M_, kTXF>), dim3(xcd_blocks((unsigned)rows * (S::L2 / S::C5))), dim3(kTXF), 0, s,
                               static_cast<const TC *>(r.a), static_cast<TC *>(r.b), tw, H, r.ms, rows);
        }
            break;
        case BIG_ZFWD_A:
            hipLaunchKernelGGL((k_zfwd_a<TC, M_, kTZ>), dim3(xcd_blocks((S::L2 / 2) * tiles)), dim3(kTZ), 0, s,
                               static_cast<const TC *>(r.a), static_cast<TC *>(r.b), tw, H, r.hs, r.ms);
            break;
        case BIG_ZFWD_B:
            if (r.uiuj)
                hipLaunchKernelGGL((k_zfwd_b<TC, M_, kTZ7U, true>), dim3(S::L1 * tiles), dim3(kTZ7U), 0, s,
                                   static_cast<const TC *>(r.a), r.om, r.fn, tw, H, r.hs, r.cn);
            else
                hipLaunchKernelGGL((k_zfwd_b<TC, M_, kTZ7>), dim3(S::L1 * tiles), dim3(kTZ7), 0, s, static_cast<const TC *>(r.a),
                                   r.om, r.fn, tw, H, r.hs, r.cn);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// natural-order inverse line FFTs for the diagnostics (FP64)
template <int M_>
hipError_t launch_lines_m(const BigReq &r)
{
    using S = Split<M_>;
    constexpr int L1 = S::L1, L2 = S::L2;
    constexpr int C = (L1 % 8 == 0 && L2 % 8 == 0) ? 8 : ((L1 % 6 == 0 && L2 % 6 == 0) ? 6 : 2);
    const cd *tw = static_cast<const cd *>(r.tw);
    hipLaunchKernelGGL((k_line_a<L1, L2, C, true>), dim3(L2 / C, (unsigned)r.lines), dim3(kT), 0, r.stream,
                       static_cast<const cd *>(r.a), static_cast<cd *>(r.scratch), tw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_line_b<L1, L2, C, true>), dim3(L1 / C, (unsigned)r.lines), dim3(kT), 0, r.stream,
                       static_cast<const cd *>(r.scratch), static_cast<cd *>(r.b), tw);
    return hipGetLastError();
}

template <class TC>
hipError_t launch_t(const BigReq &r, int M)
{
    return with_big_m(M, [&](auto m) { return launch_m<TC, decltype(m)::value>(r); });
}

// the matrix-core line transforms (cm3, cmh) on the N=29400 grid and on one size the
// oracle follows (N = 1024)
template <class TC>
hipError_t launch_t_mfma(const BigReq &r, int M)
{
    switch (M) {
        case 1536: return launch_m<TC, 1536>(r);
        case 44100: return launch_m<TC, 44100>(r);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_lines(const BigReq &r, int M)
{
    return with_big_m(M, [&](auto m) { return launch_lines_m<decltype(m)::value>(r); });
}

// partial CFL maxima written by the CFL form of pass 4: one per workgroup
template <int M_>
constexpr int cfl_parts_of() { return M_ / (2 * xb_pairs<M_, 1>()) * (Split<M_>::L1 / Split<M_>::C4); }

}  // namespace

int dns_big_supported_m(int M)
{
    return with_big_m(M, [](auto) { return hipSuccess; }) == hipSuccess ? 1 : 0;
}

int dns_big_mfma_supported_m(int M) { return (M == 1536 || M == 44100) ? 1 : 0; }

int dns_big_cfl_parts(int M)
{
    int t = 0;
    (void)with_big_m(M, [&](auto m) {
        t = cfl_parts_of<decltype(m)::value>();
        return hipSuccess;
    });
    return t;
}

hipError_t dns_big_launch(const BigReq &r)
{
    const int M = 3 * r.H;
    if (r.uiuj && r.fp32) return hipErrorInvalidValue;  // the u_i u_j form is built for FP64
    switch (r.what) {
        case BIG_LINE_INV: return launch_lines(r, M);
        case BIG_ROWS:
            hipLaunchKernelGGL(k_rows, dim3(grid_for((size_t)M * M)), dim3(256), 0, r.stream,
                               static_cast<const cd *>(r.a), static_cast<cd *>(r.b), r.H, M, r.hs, r.planes);
            return hipGetLastError();
        case BIG_ROWS_PHYS:
            hipLaunchKernelGGL(k_rows_phys, dim3(grid_for((size_t)M * M)), dim3(256), 0, r.stream,
                               static_cast<const cd *>(r.a), (size_t)M * M, r.out0, r.out_mode);
            return hipGetLastError();
        case BIG_ENERGY:
            hipLaunchKernelGGL(k_energy_tiles, dim3(dns_state_tiles(r.H)), dim3(256), 0, r.stream, r.om, 2 * r.H, r.out0,
                               r.out1);
            return hipGetLastError();
        case BIG_STATE_TILE:
        case BIG_STATE_UNTILE:
            hipLaunchKernelGGL(k_state_tile, dim3(grid_for((size_t)2 * r.H * r.H)), dim3(256), 0, r.stream,
                               static_cast<const cd *>(r.a), static_cast<cd *>(r.b), 2 * r.H,
                               r.what == BIG_STATE_TILE ? 1 : 0);
            return hipGetLastError();
        default:
            if (!r.fp32) return launch_t<cd>(r, M);
            if (r.engine == 1) return launch_t_mfma<cm3>(r, M);
            if (r.engine == 2) return launch_t_mfma<cmh>(r, M);
            return launch_t<cf>(r, M);
    }
}
