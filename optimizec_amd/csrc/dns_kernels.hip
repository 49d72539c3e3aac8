// dns_kernels.hip -- the time step of the 2D pseudo-spectral vorticity solver as
// hand-written gfx950 kernels (FP64).
//
// Grid: base N, H = N/2 retained kx (and kz = -H..H-1), M = 3N/2 = 3H physical points.
// Every M-point transform of the 3/2-padded grid is split as M = 3 x H:
//   inverse  x[z2 + 3 z1] = sum_k1 W_H^{-z1 k1} W_M^{-z2 k1} (X[k1] + X[k1+2H] W_3^{-2 z2})
//   forward  X[k1 + H k2] = sum_z2 W_M^{z2 k1} W_3^{z2 k2} F_z2[k1],
//            F_z2 = H-point FFT of x[z2 + 3 z1]
// where the zero-padded band k in [H, 2H) never enters (pruned) and only the retained
// outputs k2 in {0, 2} are formed.  Each M-point transform is therefore three
// independent H-point FFTs (LDS-resident, fft_lds.hpp) plus a radix-3 combine in
// registers.
//
// Device layouts (all FP64 complex = 16 B):
//   om, fn : [kx][iz]   column-contiguous (H columns of N)          -- state
//   V      : [f][z/P][kx][z%P] 4 fields (u, w, d_x omega, d_z omega), groups of P rows
//            (vrow)                                             -- z-inverse outputs
//   G      : [kx/Q][z][kx%Q] M x H, groups of Q modes (gpos)        -- x-forward of u.grad(omega)
//   tw     : W_M^t = e^{-2 pi i t / M}, t in [0, M)
//
// Per step two kernels run (reference step order mem_cuda.cu:1851-1857):
//   k_row  (z-row per workgroup)  : V rows -> x-inverse (u + i w, wx + i wz) -> product
//                                   P = u wx + w wz (+ CFL max) -> x-forward of P -> G row
//                                   [replaces STEP2B: step2b_cuda.cu:137-276 + cuFFT R2C]
//   k_col  (C columns per group)  : G columns -> z-forward -> FN = -P^/M^2 -> CN/AB2 update
//                                   of om, fn -> 4 spectra of the new omega -> z-inverse -> V
//                                   [replaces STEP3 + STEP2A: step3_cuda.cu:35-173,
//                                    step2a_cuda.cu:59-261 + cuFFT C2R]
#include <hip/hip_runtime.h>

#include <type_traits>

#include <cstddef>
#include <cstdint>

#include "fft_lds.hpp"
#include "dns_kernels.h"

using dnsfft::cd;
using namespace dnsfft;

namespace {

// wavenumber of compact row iz (pao_km3_cuda.cu:87-92): iz for iz < N/2, iz - N otherwise
__device__ __forceinline__ double gamma_of(int iz, int N) { return (iz < (N >> 1)) ? double(iz) : double(iz - N); }

// Spectrum of field f at (kx, kz) from omega: every field is s * omega or i s * omega
// with s real (step3_cuda.cu:102-138 for u, w; om2phys_cuda.cu:216-251 for psi).
// inv_k2 = 1 / (K^2 + 1e-30) and inv_g = 1 / gamma (0 where the reference sets 0) are
// computed once per mode by the caller, so the per-field work is multiplications only.
__device__ __forceinline__ cd field_spec(int f, cd om, int ix, int iz, double ax, double gz, double inv_k2,
                                         double inv_g)
{
    double s;
    bool times_i = true;
    switch (f) {
        case DNS_F_U:  // -i gamma / K^2; kx = 0: -i / gamma
            s = (ix >= 1) ? -gz * inv_k2 : -inv_g;
            break;
        case DNS_F_W:  // i alfa / K^2; kx = 0: 0
            s = (ix >= 1) ? ax * inv_k2 : 0.0;
            break;
        case DNS_F_OX:
            s = ax;
            break;
        case DNS_F_OZ:
            s = gz;
            break;
        case DNS_F_PSI:
            s = (ix == 0 && iz == 0) ? 0.0 : inv_k2;
            times_i = false;
            break;
        default:  // DNS_F_OMEGA
            s = 1.0;
            times_i = false;
            break;
    }
    return times_i ? mk(-s * om.y, s * om.x) : mk(s * om.x, s * om.y);
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2).  Give
// the workgroups of one XCD consecutive tiles, so that neighbouring column tiles --
// which share every 128-byte line of the row-major G and V arrays -- meet in one L2.
__device__ __forceinline__ int xcd_tile(int b, int n)
{
    return (n % 8 == 0) ? (b % 8) * (n / 8) + b / 8 : b;
}

// V layout: groups of P = vp<TC>(H) consecutive rows interleaved, [f][z / P][kx][z mod P]:
// row z of a field starts at vrow(z) and holds mode kx at + P kx.  The column kernel writes
// each mode's P rows of a group as one run of P elements; a row reads its modes P elements
// apart, so the 64 modes of one load instruction lie in 64 P elements (with the modes M apart,
// as in a column-major V, they lay 6 MB apart at N = 4096, each in its own line and page),
// and the other rows of the group, run by neighbouring workgroups of the same XCD, use the
// rest of each line.  Timed at N = 4096 (step ms, column-major V 1.20 FP64 / 0.662 FP32):
// runs of 32 / 64 / 128 bytes 1.21 / 1.14 / 1.16 ms FP64, 16 / 32 / 64 / 128 bytes
// 0.76 / 0.683 / 0.645 / 0.633 ms FP32; the FP64 ensemble sizes (H <= 512) keep 128-byte
// runs (the ensemble step with 64: 1.21 -> 1.23 ms).
// (cs, FP16 pairs, keeps the element counts of cf: runs of 64 bytes; 128-byte runs timed
// slower, row pass 0.284 -> 0.304 ms at N = 4096)
template <class TC>
__host__ __device__ constexpr int vp(int H)
{
    int p = (sizeof(TC) == 16 && H >= 1024 ? 64 : 128) / (int)sizeof(TC);
    while ((3 * H) % p != 0) p /= 2;  // whole groups: P divides M (H = 8 in FP32: P = 8)
    return p;
}
template <class TC>
__host__ __device__ constexpr size_t vrow(int z, int H)
{
    return (size_t)(z / vp<TC>(H)) * (vp<TC>(H) * (size_t)H) + (z % vp<TC>(H));
}
// G layout, the same idea across the columns: groups of Q = gq<TC>(H) consecutive modes,
// [kx / Q][z][kx mod Q]; a row writes runs of Q modes, a column reads its rows Q elements
// apart (with a row-major G, H elements apart).  Runs of 64 bytes timed best at N = 4096
// (row-major G -> 64-byte runs: step 1.148 -> 1.118 ms FP64, column kernel 0.567 -> 0.535;
// 0.640 -> 0.633 ms FP32; 128-byte runs: FP64 1.17 ms); the FP64 ensemble sizes keep
// 128-byte runs; cs keeps the element count of cf (32-byte runs; 64-byte: step 0.574 ->
// 0.579 ms).
template <class TC>
__host__ __device__ constexpr int gq(int H)
{
    int q = (sizeof(TC) == 16 && H < 1024 ? 128 : 64) / (int)sizeof(TC);
    while (H % q != 0) q /= 2;  // whole groups: Q divides H
    return q;
}

// FP16 storage of V and G (line type cs; FP32 transforms in registers and LDS).  A stored
// value is an FP16 pair h with the value h * 2^e; e is one exponent per stored line: per
// field and mode kx for V (the column kernel writes a mode's column of a field whole; the
// four fields' exponents of mode kx packed as bytes into vexp[kx]), per row z for G (gexp[z]).
// FP16 keeps a relative precision of 2^-11 at every magnitude from 2^-14 to 2^16, so the
// exponent has only to keep the line below 2^16: it comes from an upper bound on the line's
// values that the kernel already has before the transform -- the sum of |Re| + |Im| over the
// spectrum the inverse transforms (V), the sum of |P| over the physical row the forward
// transforms (G) -- scaled into [2^14, 2^15).  Values far below the line's bound lose
// relative precision, but their absolute error stays below 2^-25 of that bound.  The sums
// go through a wavefront sum and one LDS slot per wave, read after the transform (the
// barriers are the transform's own); one LDS atomic max per lane instead (a bound from
// the largest value) cost both passes 8-15%.
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int s16_exp(float bound)
{
    int E;
    (void)frexpf(bound, &E);  // bound < 2^E
    const int e = E - 15;
    return e < -110 ? -110 : (e > 110 ? 110 : e);
}
__device__ __forceinline__ unsigned s16_pack(float x, float y, int e)
{
    const h2 t{(_Float16)ldexpf(x, -e), (_Float16)ldexpf(y, -e)};
    return __builtin_bit_cast(unsigned, t);
}
__device__ __forceinline__ cs s16_unpack(unsigned raw, int e)
{
    const h2 t = __builtin_bit_cast(h2, raw);
    return cs{ldexpf((float)t.x, e), ldexpf((float)t.y, e)};
}
// exponent of field slot f (0..3) in a packed vexp word
__device__ __forceinline__ int s16_field_exp(unsigned w, int f) { return (int)(signed char)(w >> (8 * f)); }
// Wavefront sum (whole 64-lane waves), in VALU lane moves only: DPP quad permutes and row
// rotations sum each row of 16 lanes, then v_permlane32_swap and v_permlane16_swap add the
// other rows (a lane swapped with itself: the two results hold its own and its partner's
// value, in an order that does not matter for the sum).  Lanes may differ in the last bit
// (association order); callers take one lane's value.  Through __shfl_xor (six LDS
// permutes) the column pass's four field sums cost it 3%.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float wave_sum(float v)
{
    v += dpp_move<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_move<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_move<0x124>(v);  // row_ror 4
    v += dpp_move<0x128>(v);  // row_ror 8
    const auto h = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = __uint_as_float(h[0]) + __uint_as_float(h[1]);
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
template <class TC>
__host__ __device__ constexpr size_t gpos(int z, int kx, int H)
{
    return (size_t)(kx / gq<TC>(H)) * (3 * (size_t)H * gq<TC>(H)) + (size_t)z * gq<TC>(H) + (kx % gq<TC>(H));
}
// The nonlinear term's form.  Advective (u d_x omega + w d_z omega, the default): four
// z-inverted fields (u, w, d_x omega, d_z omega) in V, one real product in G.  The u_i u_j
// (divergence) form, the reference's own (step2b_cuda.cu:137-173, step3_cuda.cu:75-81):
// FN = [alpha gamma (u^2 - w^2)^ + (gamma^2 - alpha^2) (u w)^] / M^2 from two fields in V
// (u, w) and two real products in G (D = u^2 - w^2, C = u w; planes 0, 1): fewer bytes and
// line transforms per step (row: 3 inverse + 3 forward lines instead of 6 + 2; column: 6
// forward + 6 inverse instead of 3 + 12; FP64 intermediates 96 instead of 120 N^2 bytes).  Its
// wavenumber factors multiply the products' rounding by up to K^2 after the transforms (the
// advective form applies them before), which FP64 absorbs (8.8e-14 from the advective oracle
// after 100 steps at N = 4096) and FP32 transforms do not (4.8e-5, over the 1e-5 bar): built
// for FP64 at H = 2048 (N = 4096) only.
__host__ __device__ constexpr bool uiuj_built(int H) { return H == 2048; }

// the M = 3H rows group exactly
#define DNS_V_PAIRS(H_) static_assert((3 * (H_)) % vp<cd>(H_) == 0 && (3 * (H_)) % vp<dnsfft::cf>(H_) == 0, "V row groups")

__device__ __forceinline__ double inv_k2_of(double ax, double gz) { return 1.0 / (ax * ax + gz * gz + 1.0e-30); }
__device__ __forceinline__ double inv_g_of(int iz, double gz) { return (iz >= 1 && gz != 0.0) ? 1.0 / gz : 0.0; }

}  // namespace

// =====================================================================================
// Column kernel: z-direction work on C adjacent kx columns (one workgroup).
//   forward : the 3C lines z = 3 z1 + z2 of G are transformed together (H-point
//             forward FFTs), combined with W_M^{z2 k1} W_3^{z2 k2} into the retained
//             kz, FN = -P^/M^2, CN/AB2 update of om and fn
//   inverse : for each residue z2, the spectra of the nfields requested fields
//             (always 4C lines transformed; unused ones are ignored) are pre-twiddled
//             W_M^{-z2 k1} (S[k1] + S[k1+2H] W_3^{-2 z2}), inverse-transformed and written
//             to rows z = z2 + 3 z1 of V.
// =====================================================================================
// Line transforms with the opaque thread index (fft_lines OPQ) where that timed faster
// (row kernel: FP64 H = 768 0.129 -> 0.115 ms, H = 1024 0.216 -> 0.148 ms; FP32 H = 768,
// 1024, 1536 0.365 -> 0.238 ms; column kernel: FP64 H = 1024 0.128 -> 0.119 ms); slower
// or equal at the ensemble sizes.  FP64 at H = 2048: it brings both kernels under 128
// VGPRs, so they run 1024 threads per workgroup (row 0.82 -> 0.79-0.81 ms, column 0.55 ->
// 0.535 ms; at 512 threads the opaque index made them slower)
template <class TC>
constexpr bool opq_row(int H)
{
    return H == 768 || H == 1024 || H == 2048 || (!std::is_same<TC, cd>::value && H == 1536);
}
template <class TC>
constexpr bool opq_col(int H)
{
    return H == 2048 || (std::is_same<TC, cd>::value && H == 1024);
}
// In-place line transforms (fft_lines_to_pos / fft_lines_from_pos: one barrier per stage)
// where they timed faster.  The persistent FP64 row kernel (k_row_p) takes them always
// (0.748 -> 0.713 ms at N = 4096).  Timed slower or equal elsewhere on one box, so off: the
// other row kernels' inverse + forward pairs (FP32 row kernel at N = 4096 0.345 -> 0.351 ms,
// ensemble step 1.23 -> 1.27 ms) and the column kernels' forward lines, whose G rows are then
// gathered in position order (FP64 0.499 -> 0.512 ms, FP32 0.319 -> 0.329 ms).
template <class TC>
constexpr bool ip_row(int H)
{
    return false;
}
template <class TC>
constexpr bool ip_col(int H)
{
    return false;
}
// PLAN (the ensemble autotuner's launch variants): bit 0 forces the opaque FFT thread index,
// bit 1 the radix-16 plan of the line transforms (different rounding: not bitwise equal)
template <class TC, int H, int C, int T, int MINW, int PLAN = 0, bool DIV = false>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(MINW))) void k_col(cd *__restrict__ om, cd *__restrict__ fn,
                                           const TC *__restrict__ G, TC *__restrict__ V,
                                           const TC *__restrict__ twT, const cd *__restrict__ tw, const ColArgs a,
                                           BatchArgs bt, const TC *__restrict__ twH,
                                           const int *__restrict__ gexp, unsigned *__restrict__ vexp)
{
    DNS_V_PAIRS(H);
    constexpr int N = 2 * H;
    constexpr int M = 3 * H;
    constexpr int PJ = C * H / T;
    // cs: V and G stored as FP16 pairs (exponents gexp per G row, vexp per V column); one
    // column per workgroup, unbatched
    constexpr bool S16 = std::is_same<TC, cs>::value;
    static_assert(!S16 || C == 1, "FP16 storage: one column per workgroup");
    // CN coefficients: the member's own under batching.  `a` itself is never written, so
    // a.fields[fi] (a run-time index) is read from the kernel arguments, not from a
    // private copy of the struct in scratch
    double VT = a.VT, c2 = a.c2, c3 = a.c3;
    {
        const size_t b = blockIdx.y;
        om += b * bt.s_om;
        fn += b * bt.s_om;
        if (G) G += b * bt.s_g;
        V += b * bt.s_v;
        if (bt.coef) {
            VT = bt.coef[3 * b];
            c2 = bt.coef[3 * b + 1];
            c3 = bt.coef[3 * b + 2];
        }
    }
    constexpr int LS = padded_len(H);
    static_assert((C * H) % T == 0, "tile must divide evenly over the threads");
    __shared__ TC lds[3 * C * LS];
    // Small H (the ensemble sizes): every twiddle from LDS -- W_H^j in the transform type for
    // the H-point lines, and in FP64 for the residue factors W_M^e = W_H^{e / 3} W_M^{e mod 3}
    // -- so the om, fn loads of the CN update can be issued before the forward transforms
    // without any load inside them waiting behind those (vector loads complete in order).
    constexpr bool PF = H <= 512;
    constexpr bool F64 = std::is_same<TC, cd>::value;
    // TAB: the twiddle tables in LDS -- at the small H, and for FP64 at H = 2048, where the
    // table fits beside the three lines without costing a workgroup per CU (one fits anyway;
    // with the field loop's opaque index below, 0.534 -> 0.525 ms)
    constexpr bool TAB = PF || (F64 && H == 2048);
    // TABL: the H-point lines' table W_H^j in LDS; FP32 at H = 2048 takes it alone (16 KB:
    // two workgroups per CU still fit; 0.326 -> 0.321 ms), its residue factors stay global
    constexpr bool TABL = TAB || (!F64 && H == 2048);
    __shared__ TC twl[TABL ? H : 1];
    __shared__ cd twd[TAB && !F64 ? H : 1];

    const int tid = threadIdx.x;
    const int kx0 = xcd_tile(blockIdx.x, gridDim.x) * C;
    const cd wm1 = tw[1], wm2 = tw[2];
    if constexpr (TABL) {
        for (int j = tid; j < H; j += T) {
            twl[j] = twH ? twH[j] : twT[3 * j];
            if constexpr (!F64 && TAB) twd[j] = tw[3 * j];
        }
        __syncthreads();
    }
    const cd *tabd = nullptr;
    if constexpr (TAB) tabd = F64 ? reinterpret_cast<const cd *>(twl) : twd;
    auto wM = [&](int e) -> cd {  // W_M^e, 0 <= e < 2H
        if constexpr (TAB) {
            const int q3 = e / 3, r3 = e - 3 * q3;
            const cd t = tabd[q3];
            const cd m = mul(t, cd{r3 == 1 ? wm1.x : wm2.x, r3 == 1 ? wm1.y : wm2.y});
            return cd{r3 == 0 ? t.x : m.x, r3 == 0 ? t.y : m.y};
        } else {
            return tw[opaque(e)];
        }
    };
    constexpr bool OQ = opq_col<TC>(H) || (PLAN & 1) != 0;
    constexpr int MAXR = (PLAN & 2) ? 16 : 8;
    auto lines = [&](auto inv) {  // the 3C H-point lines in LDS
        if constexpr (TABL)
            fft_lines<H, LS, 3 * C, T, decltype(inv)::value, MAXR, OQ>(lds, twl, 1);
        else
            fft_lines<H, LS, 3 * C, T, decltype(inv)::value, MAXR, OQ>(lds, twT, 3);
    };
    // the forward lines take their input in position order (in place, see fft_lines_from_pos):
    // position p of a residue line holds z1 = zpos(p); G is read one value per row anyway
    auto lines_fwd = [&]() {
        if constexpr (TABL)
            fft_lines_from_pos<H, LS, 3 * C, T, false, MAXR, OQ, ip_col<TC>(H)>(lds, twl, 1);
        else
            fft_lines_from_pos<H, LS, 3 * C, T, false, MAXR, OQ, ip_col<TC>(H)>(lds, twT, 3);
    };
    auto zpos = [](int p) { return lines_in_place<TC, ip_col<TC>(H)>() ? ip_idx(H, p, MAXR) : p; };

    // registers: the (lo, hi) pair of each owned (column c, k1): iz = k1 and iz = k1 + H
    cd xlo[PJ], xhi[PJ];

    // u_i u_j form (DIV): G holds D^ (plane 0) and C^ (plane 1); FN = (alpha gamma D^ +
    // (gamma^2 - alpha^2) C^) / M^2, accumulated in xlo / xhi plane by plane
    static_assert(!DIV || (uiuj_built(H) && std::is_same<TC, cd>::value), "u_i u_j form: FP64 at H = 2048");
    constexpr int NPL = DIV ? 2 : 1;
    if (a.do_forward) {
        // om, fn of the owned modes (lo: iz = k1, hi: iz = k1 + H), in flight during the
        // forward transforms when the twiddles come from LDS
        cd olo[PJ], ohi[PJ], flo[PJ], fhi[PJ];
        // u_i u_j form: plane 0's share of FN, held across plane 1's transform
        cd plo[DIV ? PJ : 1], phi[DIV ? PJ : 1];
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl) {
            if (pl > 0) __syncthreads();  // the previous plane's combine has read the lines
            for (int e = tid; e < 3 * C * H; e += T) {
                const int c = e % C, q = e / C;
                const int p = q / 3, z2 = q - 3 * p;
                const int z = 3 * zpos(p) + z2;
                const size_t go = (size_t)pl * a.vstride + gpos<TC>(z, kx0 + c, H);
                if constexpr (S16)
                    lds[(z2 * C + c) * LS + pidx(p)] = s16_unpack(reinterpret_cast<const unsigned *>(G)[go], gexp[z]);
                else
                    lds[(z2 * C + c) * LS + pidx(p)] = G[go];
            }
            if constexpr (PF && !DIV) {
#pragma unroll
                for (int j = 0; j < PJ; ++j) {
                    const int p = tid + j * T;
                    const int c = p / H, k1 = p - (p / H) * H;
                    const size_t io = (size_t)(kx0 + c) * N + k1;
                    olo[j] = om[io];
                    ohi[j] = om[io + H];
                    flo[j] = fn[io];
                    fhi[j] = fn[io + H];
                }
            }
            __syncthreads();
            lines_fwd();
            // radix-3 combine in FP64 (the FFT outputs are widened)
            const cd w3a = tw[2 * H];  // W_3^2
            const cd w3b = tw[H];      // W_3^4 = W_3
#pragma unroll
            for (int j = 0; j < PJ; ++j) {
                const int p = tid + j * T;
                const int c = p / H, k1 = p - (p / H) * H;
                const cd F0 = to_d(lds[(0 * C + c) * LS + pidx(k1)]);
                const cd F1 = mul(to_d(lds[(1 * C + c) * LS + pidx(k1)]), wM(k1));
                const cd F2 = mul(to_d(lds[(2 * C + c) * LS + pidx(k1)]), wM(2 * k1));
                const cd Xlo = add(F0, add(F1, F2));
                const cd Xhi = add(F0, add(mul(F1, w3a), mul(F2, w3b)));
                if constexpr (DIV) {
                    const double ax = double(kx0 + c), glo = gamma_of(k1, N), ghi = gamma_of(k1 + H, N);
                    if (pl == 0) {
                        plo[j] = scl(Xlo, glo * ax * a.divxz);
                        phi[j] = scl(Xhi, ghi * ax * a.divxz);
                    } else {
                        xlo[j] = add(plo[j], scl(Xlo, (glo * glo - ax * ax) * a.divxz));
                        xhi[j] = add(phi[j], scl(Xhi, (ghi * ghi - ax * ax) * a.divxz));
                    }
                } else {
                    xlo[j] = Xlo;
                    xhi[j] = Xhi;
                }
            }
        }
        // CN/AB2 update (step3_cuda.cu:75-100): FN = -P^/M^2 (advective; the u_i u_j form has
        // FN in xlo / xhi already), row kz = -N/2 forced to 0
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            const int p = tid + j * T;
            const int c = p / H, k1 = p - (p / H) * H;
            const int kx = kx0 + c;
            const double ax = double(kx);
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const int iz = k1 + half * H;
                const double gz = gamma_of(iz, N);
                const double K2 = ax * ax + gz * gz;
                const cd X = half ? xhi[j] : xlo[j];
                cd FN = DIV ? X : mk(-X.x * a.divxz, -X.y * a.divxz);
                const size_t io = (size_t)kx * N + iz;
                if (a.fext) FN = add(FN, a.fext[io]);
                if (iz == H) FN = mk(0.0, 0.0);
                cd o, f;
                if constexpr (PF && !DIV) {
                    o = half ? ohi[j] : olo[j];
                    f = half ? fhi[j] : flo[j];
                } else {
                    o = om[io];
                    f = fn[io];
                }
                const double ARG = VT * K2;
                const double DEN = 1.0 + ARG;
                const double c1 = 1.0 - ARG;
                const cd num = mk(c1 * o.x + c2 * FN.x + c3 * f.x, c1 * o.y + c2 * FN.y + c3 * f.y);
                const cd on = mk(num.x / DEN, num.y / DEN);
                om[io] = on;
                fn[io] = FN;
                if (half)
                    xhi[j] = on;
                else
                    xlo[j] = on;
            }
        }
        __syncthreads();
    } else {
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            const int p = tid + j * T;
            const int c = p / H, k1 = p - (p / H) * H;
            const size_t base = (size_t)(kx0 + c) * N;
            xlo[j] = om[base + k1];
            xhi[j] = om[base + k1 + H];
        }
    }

    // per-mode reciprocals shared by all fields
    double ik_lo[PJ], ik_hi[PJ];
#pragma unroll
    for (int j = 0; j < PJ; ++j) {
        const int p = tid + j * T;
        const int c = p / H, k1 = p - (p / H) * H;
        const double ax = double(kx0 + c);
        ik_lo[j] = inv_k2_of(ax, gamma_of(k1, N));
        ik_hi[j] = inv_k2_of(ax, gamma_of(k1 + H, N));
    }

    // z-inverse of each requested spectrum of the (new) omega: the three residues z2
    // are transformed together, then each column of V (column-major, z contiguous) is
    // written whole
    const int nf = a.nfields;
    __shared__ float redf[S16 ? T / 64 : 1];  // cs: per-wave partial bounds of a field's column
    unsigned vw = 0;                          // cs: the column's packed field exponents
#pragma unroll 1
    for (int fi = 0; fi < nf; ++fi) {
        const int f = a.fields[fi];
        // this field's own copy of the thread index where the transforms take the opaque one
        // (FP32, and FP64 at H = 2048), so its index arithmetic stays in the loop instead of
        // being hoisted into registers (FP32 at N = 4096: 60 VGPRs and no spill, column
        // kernel 0.37 -> 0.33 ms; FP64 at H = 1024 is not faster with it)
        constexpr bool FOPQ = (opq_col<TC>(H) && (!std::is_same<TC, cd>::value || H == 2048)) || (PLAN & 1) != 0;
        const int tid = FOPQ ? opaque((int)threadIdx.x) : (int)threadIdx.x;
        double l1 = 0.0;  // cs: sum of |Re| + |Im| over this thread's spectrum values
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            const int p = tid + j * T;
            const int c = p / H, k1 = p - (p / H) * H;
            const int kx = kx0 + c;
            const double ax = double(kx);
            const double glo = gamma_of(k1, N), ghi = gamma_of(k1 + H, N);
            // 1/gamma is only needed on the kx = 0 column
            const double iglo = (kx == 0) ? inv_g_of(k1, glo) : 0.0;
            const double ighi = (kx == 0) ? inv_g_of(k1 + H, ghi) : 0.0;
            const cd slo = field_spec(f, xlo[j], kx, k1, ax, glo, ik_lo[j], iglo);
            const cd shi = field_spec(f, xhi[j], kx, k1 + H, ax, ghi, ik_hi[j], ighi);
            // residue z2: W_M^{-z2 k1} (S[k1] + S[k1 + 2H] W_3^{-2 z2})
            lds[(0 * C + c) * LS + pidx(k1)] = from_d<TC>(add(slo, shi));
            lds[(1 * C + c) * LS + pidx(k1)] = from_d<TC>(mulc(add(slo, mulc(shi, tw[2 * H])), wM(k1)));
            lds[(2 * C + c) * LS + pidx(k1)] = from_d<TC>(mulc(add(slo, mulc(shi, tw[H])), wM(2 * k1)));
            if constexpr (S16) l1 += fabs(slo.x) + fabs(slo.y) + fabs(shi.x) + fabs(shi.y);
        }
        if constexpr (S16) {
            const float w = wave_sum((float)l1);
            if ((threadIdx.x & 63) == 0) redf[threadIdx.x >> 6] = w;
        }
        __syncthreads();
        // cs: the column's values are bounded by the sum of |Re| + |Im| of its spectrum (read
        // before the transform, which hides the reads' latency)
        int ex = 0;
        if constexpr (S16) {
            float bound = 0.0f;
#pragma unroll
            for (int w = 0; w < T / 64; ++w) bound += redf[w];
            ex = __builtin_amdgcn_readfirstlane(s16_exp(bound));  // the same in every lane: a scalar
        }
        lines(std::true_type{});
        if constexpr (S16) {
            unsigned *out = reinterpret_cast<unsigned *>(V) + (size_t)fi * a.vstride;
            for (int e = tid; e < M; e += T) {
                const int z1 = e / 3, z2 = e - 3 * z1;
                const TC v = lds[z2 * LS + pidx(z1)];
                out[vrow<TC>(e, H) + vp<TC>(H) * (size_t)kx0] = s16_pack(v.x, v.y, ex);
            }
            vw |= (unsigned)(ex & 0xff) << (8 * fi);
        } else {
            TC *out = V + (size_t)fi * a.vstride;
            for (int e = tid; e < C * M; e += T) {
                const int c = e / M, zz = e - c * M;
                const int z1 = zz / 3, z2 = zz - 3 * z1;
                out[vrow<TC>(zz, H) + vp<TC>(H) * (size_t)(kx0 + c)] = lds[(z2 * C + c) * LS + pidx(z1)];
            }
        }
        __syncthreads();
    }
    if constexpr (S16) {
        if (threadIdx.x == 0) vexp[kx0] = vw;
    }
}

// =====================================================================================
// Persistent column kernel (the step's column pass at H = 2048, FP64): CPW columns per
// workgroup, one workgroup per CU.  Per column the work is k_col's line for line (so the
// results are bitwise equal); the twiddle table is filled once per workgroup and the next
// column's G values are gathered while the current column's last inverse transform runs.
// =====================================================================================
template <int H, int T, int CPW>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(1))) void k_col_p(
    cd *__restrict__ om, cd *__restrict__ fn, const cd *__restrict__ G, cd *__restrict__ V, const cd *__restrict__ twT,
    const cd *__restrict__ tw, const ColArgs a, const cd *__restrict__ twH)
{
    DNS_V_PAIRS(H);
    using TC = cd;
    constexpr int N = 2 * H;
    constexpr int M = 3 * H;
    constexpr int PJ = H / T;
    constexpr int GJ = 3 * H / T;  // G values per thread and column
    static_assert(H % T == 0, "whole half columns per thread");
    const double VT = a.VT, c2 = a.c2, c3 = a.c3;
    constexpr int LS = padded_len(H);
    __shared__ TC lds[3 * LS];
    __shared__ TC twl[H];
    const int nw = gridDim.x;  // a multiple of 8 with CPW * nw = H (checked by the launcher)
    auto col_of = [&](int i) {
        const int b = blockIdx.x, x = b & 7, c = b >> 3;
        return x * (H / 8) + i * (nw / 8) + c;
    };
    const cd wm1 = tw[1], wm2 = tw[2];
    for (int j = threadIdx.x; j < H; j += T) twl[j] = twH ? twH[j] : twT[3 * j];
    __syncthreads();
    auto wM = [&](int e) -> cd {
        const int q3 = e / 3, r3 = e - 3 * q3;
        const cd t = twl[q3];
        const cd m = mul(t, cd{r3 == 1 ? wm1.x : wm2.x, r3 == 1 ? wm1.y : wm2.y});
        return cd{r3 == 0 ? t.x : m.x, r3 == 0 ? t.y : m.y};
    };
    auto lines = [&](auto inv) { fft_lines<H, LS, 3, T, decltype(inv)::value, 8, true>(lds, twl, 1); };
    // the forward lines take their input in position order (in place): slot q = 3 p + z2 of
    // the gather holds G row 3 ip_idx(p) + z2
    auto grow = [](int q) {
        const int p = q / 3, z2 = q - 3 * p;
        return 3 * (ip_col<cd>(H) ? ip_idx(H, p) : p) + z2;
    };
    // the G values of a column (slots q = tid + r T), in registers: copied component-wise,
    // since a whole-struct copy between global memory and this array becomes a memcpy that
    // keeps the array in scratch (and makes the prefetch wait for its loads)
    cd gx[GJ];
#pragma unroll
    for (int r = 0; r < GJ; ++r) {
        const cd v = G[gpos<cd>(grow(threadIdx.x + r * T), col_of(0), H)];
        gx[r] = mk(v.x, v.y);
    }
#pragma unroll 1
    for (int i = 0; i < CPW; ++i) {
        const int tid = opaque((int)threadIdx.x);
        const int kx0 = col_of(i);
        cd xlo[PJ], xhi[PJ];
#pragma unroll
        for (int r = 0; r < GJ; ++r) {
            const int z = tid + r * T;
            const int z1 = z / 3, z2 = z - 3 * z1;
            lds[z2 * LS + pidx(z1)] = mk(gx[r].x, gx[r].y);
        }
        __syncthreads();
        fft_lines_from_pos<H, LS, 3, T, false, 8, true, ip_col<cd>(H)>(lds, twl, 1);
        const cd w3a = tw[2 * H];
        const cd w3b = tw[H];
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            const int k1 = tid + j * T;
            const cd F0 = lds[pidx(k1)];
            const cd F1 = mul(lds[LS + pidx(k1)], wM(k1));
            const cd F2 = mul(lds[2 * LS + pidx(k1)], wM(2 * k1));
            xlo[j] = add(F0, add(F1, F2));
            xhi[j] = add(F0, add(mul(F1, w3a), mul(F2, w3b)));
        }
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            const int k1 = tid + j * T;
            const int kx = kx0;
            const double ax = double(kx);
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const int iz = k1 + half * H;
                const double gz = gamma_of(iz, N);
                const double K2 = ax * ax + gz * gz;
                const cd X = half ? xhi[j] : xlo[j];
                cd FN = mk(-X.x * a.divxz, -X.y * a.divxz);
                const size_t io = (size_t)kx * N + iz;
                if (a.fext) FN = add(FN, a.fext[io]);
                if (iz == H) FN = mk(0.0, 0.0);
                const cd o = om[io];
                const cd f = fn[io];
                const double ARG = VT * K2;
                const double DEN = 1.0 + ARG;
                const double c1 = 1.0 - ARG;
                const cd num = mk(c1 * o.x + c2 * FN.x + c3 * f.x, c1 * o.y + c2 * FN.y + c3 * f.y);
                const cd on = mk(num.x / DEN, num.y / DEN);
                om[io] = on;
                fn[io] = FN;
                if (half)
                    xhi[j] = on;
                else
                    xlo[j] = on;
            }
        }
        __syncthreads();
        double ik_lo[PJ], ik_hi[PJ];
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            const int k1 = tid + j * T;
            const double ax = double(kx0);
            ik_lo[j] = inv_k2_of(ax, gamma_of(k1, N));
            ik_hi[j] = inv_k2_of(ax, gamma_of(k1 + H, N));
        }
        const int nf = a.nfields;
#pragma unroll 1
        for (int fi = 0; fi < nf; ++fi) {
            const int f = a.fields[fi];
            const int tid = opaque((int)threadIdx.x);
#pragma unroll
            for (int j = 0; j < PJ; ++j) {
                const int k1 = tid + j * T;
                const int kx = kx0;
                const double ax = double(kx);
                const double glo = gamma_of(k1, N), ghi = gamma_of(k1 + H, N);
                const double iglo = (kx == 0) ? inv_g_of(k1, glo) : 0.0;
                const double ighi = (kx == 0) ? inv_g_of(k1 + H, ghi) : 0.0;
                const cd slo = field_spec(f, xlo[j], kx, k1, ax, glo, ik_lo[j], iglo);
                const cd shi = field_spec(f, xhi[j], kx, k1 + H, ax, ghi, ik_hi[j], ighi);
                lds[pidx(k1)] = add(slo, shi);
                lds[LS + pidx(k1)] = mulc(add(slo, mulc(shi, tw[2 * H])), wM(k1));
                lds[2 * LS + pidx(k1)] = mulc(add(slo, mulc(shi, tw[H])), wM(2 * k1));
            }
            // the next column's G values, in flight during this column's last transform
            if (fi == nf - 1 && i + 1 < CPW) {
                const int kn = col_of(i + 1);
#pragma unroll
                for (int r = 0; r < GJ; ++r) {
                    const cd v = G[gpos<cd>(grow(tid + r * T), kn, H)];
                    gx[r] = mk(v.x, v.y);
                }
            }
            __syncthreads();
            lines(std::true_type{});
            TC *out = V + (size_t)fi * a.vstride + vp<TC>(H) * (size_t)kx0;
            for (int e = tid; e < M; e += T) {
                const int z1 = e / 3, z2 = e - 3 * z1;
                out[vrow<TC>(e, H)] = lds[z2 * LS + pidx(z1)];
            }
            __syncthreads();
        }
    }
}

// =====================================================================================
// Row kernel: x-direction work on one z-row of the padded grid.
// MODE 0: full nonlinear row (inverse of u+iw and wx+iwz, product, forward); the step's
//         row pass, no CFL (the step's dt comes from NEXTDT).
// MODE 1: velocity only (inverse of u+iw, CFL max) -- NEXTDT (dns_cuda.cu:485-547).
//
// The x-inverse of two real fields a, b packs Z = a + i b on the full M-point
// spectrum: Z[k] = a[k] + i b[k] (k < H; bin 0 keeps only real parts, the C2R
// convention) and Z[M - k] = conj(a[k]) + i conj(b[k]).  With M = 3H and
// x = n2 + 3 n1, each residue n2 is one H-point inverse of
//   Y[k1] = W_M^{-n2 k1} (Zlo[k1] + Zhi[k1] W_3^{-2 n2}),  Zlo = Z[k1], Zhi = Z[k1 + 2H].
// Zhi[k1] needs the input at H - k1: the input row is read once, lanes l and l^32 of a
// wave hold the pair (q, H - q), and one cross-lane exchange gives each its mirror.
// The three residues of one field pair are transformed together (3 lines).
// =====================================================================================
template <class TC, int H, int T, int MODE, int MINW, int PLAN = 0>
__global__ __launch_bounds__(T, MINW) void k_row(const TC *__restrict__ V, size_t vstride, TC *__restrict__ G,
                                           const TC *__restrict__ tw, double *__restrict__ cflrow,
                                           double DX, double DZ, BatchArgs bt, const TC *__restrict__ twH,
                                           const unsigned *__restrict__ vexp, int *__restrict__ gexp)
{
    DNS_V_PAIRS(H);
    // cs: V and G stored as FP16 pairs (exponents vexp per V column, gexp per G row); the
    // step's row pass only, unbatched
    constexpr bool S16 = std::is_same<TC, cs>::value;
    static_assert(!S16 || MODE == 0, "FP16 storage: the step's row pass");
    {
        const size_t b = blockIdx.y;
        V += b * bt.s_v;
        if (G) G += b * bt.s_g;
        if (cflrow) cflrow += b * bt.s_cfl;
    }
    using RT = decltype(TC::x);
    constexpr int LS = padded_len(H);
    constexpr int PJ = (H / T) > 0 ? (H / T) : 1;  // input pairs per lane
    constexpr int PO = (H + T - 1) / T;            // outputs per thread and residue
    static_assert(T % 64 == 0, "whole wavefronts");
    static_assert(H % T == 0 || H < T, "pair slots must tile the half row");
    __shared__ TC lds[3 * LS];
    __shared__ TC twl[H];  // W_H^j = W_M^{3j}; W_M^e = W_H^{e / 3} W_M^{e % 3}
    __shared__ double red[T / 64];  // MODE 1: per-wave CFL maxima; cs: per-wave |P| sums
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // rows in XCD-aware order: neighbouring rows share the 128-byte lines of the
    // column-major V, so they should meet in one L2
    const int z = xcd_tile(blockIdx.x, gridDim.x);
    const TC *row = V + vrow<TC>(z, H);  // element (f, k) at f * vstride + vp k
    // W_3^{2 n2} for n2 = 0, 1, 2 (indices (2 n2 H) mod M)
    const TC w3_1 = tw[2 * H], w3_2 = tw[H];
    const TC zero{RT(0), RT(0)};
    // every twiddle of the row comes from LDS: W_H^j for the H-point lines and
    // W_M^e = W_H^{e / 3} W_M^{e mod 3} for the residue pre- and post-twiddles (one extra
    // complex multiply), so no global load inside the transforms waits behind the
    // prefetched inputs (vector memory loads complete in order)
    const TC wm1 = tw[1], wm2 = tw[2];
    // (contiguous W_H table when given: one 16-byte line share per entry instead of three)
    for (int j = tid; j < H; j += T) twl[j] = twH ? twH[j] : tw[3 * j];
    __syncthreads();
    auto wM = [&](int e) {  // component-wise selects (struct selects go through scratch)
        const int q3 = e / 3, r3 = e - 3 * q3;
        const TC t = twl[q3];
        const TC m = mul(t, TC{r3 == 1 ? wm1.x : wm2.x, r3 == 1 ? wm1.y : wm2.y});
        return TC{r3 == 0 ? t.x : m.x, r3 == 0 ? t.y : m.y};
    };

    TC vel[3][PO];     // u + i w at x = n2 + 3 x1
    double cmax = 0.0;

    constexpr int NPH = (MODE == 0) ? 2 : 1;
    // raw inputs of the two field pairs: the second pair's loads are issued before the
    // first pair's transforms (whose twiddles come from LDS), so they are in flight
    // while those run.  Lane l (< 32) of a wave takes mode k = q, lane l ^ 32 mode H - q.
    // (cs: the raw FP16 pairs, unpacked where they are used, so the loads stay in flight; the
    // packed exponents of each lane's modes)
    using RAW = typename std::conditional<S16, unsigned, TC>::type;
    RAW ra0[PJ], rb0[PJ], ra1[PJ], rb1[PJ];
    unsigned vx[S16 ? PJ : 1];
    auto issue = [&](int f0, RAW *a, RAW *b) {
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            const int q = (tid >> 6) * 32 + (lane & 31) + j * (T / 2);
            const int k = lane < 32 ? q : (q == 0 ? H / 2 : H - q);
            if constexpr (S16) {
                const unsigned *fa = reinterpret_cast<const unsigned *>(V) + f0 * vstride + vrow<TC>(z, H);
                unsigned va = 0u, vb = 0u, w = 0u;
                if (q < H / 2) {
                    va = fa[vp<TC>(H) * (size_t)k];
                    vb = fa[vstride + vp<TC>(H) * (size_t)k];
                    if (f0 == 0) w = vexp[k];
                }
                if (f0 == 0) vx[j] = w;
                a[j] = va;
                b[j] = vb;
            } else {
                const TC *fa = row + f0 * vstride;
                TC va = zero, vb = zero;  // a branch, not a select (a select of the two sources
                if (q < H / 2) {          // becomes a flat load from a scratch copy of zero)
                    va = fa[vp<TC>(H) * (size_t)k];
                    vb = fa[vstride + vp<TC>(H) * (size_t)k];
                }
                a[j] = va;
                b[j] = vb;
            }
        }
    };
    auto value = [&](RAW r, int f, int j) -> TC {
        if constexpr (S16)
            return s16_unpack(r, s16_field_exp(vx[j], f));
        else
            return r;
    };
    issue(0, ra0, rb0);
#pragma unroll
    for (int ph = 0; ph < NPH; ++ph) {
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            const int q = (tid >> 6) * 32 + (lane & 31) + j * (T / 2);
            const bool valid = q < H / 2;
            const bool hi = lane >= 32;
            const int k = !hi ? q : (q == 0 ? H / 2 : H - q);
            const TC av = value(ph == 0 ? ra0[j] : ra1[j], 2 * ph, j);
            const TC bv = value(ph == 0 ? rb0[j] : rb1[j], 2 * ph + 1, j);
            // partner lane (l ^ 32) holds H - k, except in the q = 0 pair (0 and H/2)
            TC am{swap_halves(av.x), swap_halves(av.y)};
            TC bm{swap_halves(bv.x), swap_halves(bv.y)};
            if (k == H / 2) {  // its own mirror
                am = av;
                bm = bv;
            }
            if (valid) {
                TC zlo, zhi;
                if (k == 0) {
                    zlo = TC{av.x, bv.x};
                    zhi = zero;
                } else {
                    zlo = TC{av.x - bv.y, av.y + bv.x};
                    zhi = TC{am.x + bm.y, bm.x - am.y};
                }
                lds[pidx(k)] = add(zlo, zhi);
                lds[LS + pidx(k)] = mulc(add(zlo, mulc(zhi, w3_1)), wM(k));
                lds[2 * LS + pidx(k)] = mulc(add(zlo, mulc(zhi, w3_2)), wM(2 * k));
            }
        }
        if constexpr (NPH == 2) {
            if (ph == 0) issue(2, ra1, rb1);
        }
        __syncthreads();
        // vector lines in place: position p holds x1 = ip_idx(p) (the product below is pointwise
        // and the forward takes that order back); matrix-core lines in natural order
        fft_lines_to_pos<H, LS, 3, T, true, (PLAN & 2) ? 16 : 8, opq_row<TC>(H) || (PLAN & 1) != 0, ip_row<TC>(H)>(lds, twl, 1);
        float l1 = 0.0f;  // cs: sum of |P| over this thread's points of the row
#pragma unroll
        for (int j = 0; j < PO; ++j) {
            const int x1 = tid + j * T;
            if (x1 < H) {
                if (ph == 0) {
#pragma unroll
                    for (int n2 = 0; n2 < 3; ++n2) {
                        const TC v = lds[n2 * LS + pidx(x1)];
                        vel[n2][j] = v;
                        if constexpr (MODE == 1) {
                            const double cfl = fabs(double(v.x)) / DX + fabs(double(v.y)) / DZ;
                            cmax = cfl > cmax ? cfl : cmax;
                        }
                    }
                } else {
                    // the real product P = u d_x omega + w d_z omega, written straight into
                    // the x-forward's lines A = P0 + i P1 (line 0) and B = P2 (line 1): this
                    // thread alone reads and writes position x1 of the lines
                    RT P[3];
#pragma unroll
                    for (int n2 = 0; n2 < 3; ++n2) {
                        const TC v = lds[n2 * LS + pidx(x1)];
                        P[n2] = vel[n2][j].x * v.x + vel[n2][j].y * v.y;
                    }
                    lds[pidx(x1)] = TC{P[0], P[1]};
                    lds[LS + pidx(x1)] = TC{P[2], RT(0)};
                    if constexpr (S16) l1 += fabsf(P[0]) + fabsf(P[1]) + fabsf(P[2]);
                }
            }
        }
        if constexpr (S16) {
            if (ph == NPH - 1) {
                const float w = wave_sum(l1);
                if (lane == 0) red[tid >> 6] = w;
            }
        }
        __syncthreads();
    }

    if (MODE == 0) {
        // x-forward of the real product: F_n2 = H-point FFT of P[n2 + 3 x1] (lines
        // A = P0 + i P1, B = P2), then X[k] = F0 + W_M^k F1 + W_M^{2k} F2 for k < H
        TC *A = lds;
        TC *B = lds + LS;
        // cs: the row's spectrum is bounded by the sum of |P| over the row (read before the
        // transform, which hides the reads' latency)
        int ex = 0;
        if constexpr (S16) {
            float bound = 0.0f;
#pragma unroll
            for (int w = 0; w < T / 64; ++w) bound += (float)red[w];
            ex = __builtin_amdgcn_readfirstlane(s16_exp(bound));  // the same in every lane: a scalar
            if (tid == 0) gexp[z] = ex;
        }
        fft_lines_from_pos<H, LS, 2, T, false, (PLAN & 2) ? 16 : 8, opq_row<TC>(H) || (PLAN & 1) != 0, ip_row<TC>(H)>(lds, twl, 1);
        const RT half = RT(0.5);
        for (int k = tid; k < H; k += T) {
            const TC za = A[pidx(k)];
            const TC zm = conj(A[pidx(k == 0 ? 0 : H - k)]);
            const TC s = add(za, zm), d = sub(za, zm);
            const TC F0{half * s.x, half * s.y};
            const TC F1{half * d.y, -half * d.x};  // d / (2 i)
            const TC F2 = B[pidx(k)];
            const TC g = add(F0, add(mul(F1, wM(k)), mul(F2, wM(2 * k))));
            if constexpr (S16)
                reinterpret_cast<unsigned *>(G)[gpos<TC>(z, k, H)] = s16_pack(g.x, g.y, ex);
            else
                G[gpos<TC>(z, k, H)] = g;
        }
    }

    if (MODE == 1 && cflrow) {
        // wavefront max, then across the workgroup's waves
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(cmax, off, 64);
            cmax = o > cmax ? o : cmax;
        }
        if (lane == 0) red[tid >> 6] = cmax;
        __syncthreads();
        if (tid == 0) {
            double m = red[0];
            for (int i = 1; i < T / 64; ++i) m = red[i] > m ? red[i] : m;
            cflrow[z] = m;
        }
    }
}

// =====================================================================================
// Persistent row kernel (the step's MODE 0 at H = 2048): RPW rows per workgroup, one
// workgroup per CU.  Per row the work is k_row's, but the line transforms run in place
// (fft_lines_ip: one barrier per stage instead of two; a DIF inverse leaves x1 = ip_idx(p) at
// position p, the product is formed position by position and the DIT forward returns
// natural order, so nothing is permuted), the W_H table is filled once per workgroup and the
// next row's first field pair is loaded while the current row's forward transform runs.  Rows keep
// the XCD-aware order: workgroup b runs on XCD b mod 8 and takes rows x (M / 8) + i (NW / 8)
// + b / 8, so at any time the 32 workgroups of an XCD work on consecutive rows that share
// the 128-byte lines of the column-major V in that XCD's L2.
// =====================================================================================
template <class TC, int H, int T, int RPW, bool DIV = false>
__global__ __launch_bounds__(T, 1) void k_row_p(const TC *__restrict__ V, size_t vstride, TC *__restrict__ G,
                                                const TC *__restrict__ tw, const TC *__restrict__ twH)
{
    DNS_V_PAIRS(H);
    using RT = decltype(TC::x);
    constexpr int LS = padded_len(H);
    constexpr int PJ = H / T;  // input pairs per lane
    constexpr int PO = H / T;  // outputs per thread and residue
    static_assert(T % 64 == 0 && H % T == 0, "whole wavefronts; pair slots tile the half row");
    __shared__ TC lds[3 * LS];
    __shared__ TC twl[H];
    constexpr int M = 3 * H;
    int tid = threadIdx.x;
    int lane = tid & 63;
    const int nw = gridDim.x;  // a multiple of 8 with RPW * nw = M (checked by the launcher)
    auto row_of = [&](int i) {
        const int b = blockIdx.x, x = b & 7, c = b >> 3;
        return x * (M / 8) + i * (nw / 8) + c;
    };
    const TC w3_1 = tw[2 * H], w3_2 = tw[H];
    const TC zero{RT(0), RT(0)};
    const TC wm1 = tw[1], wm2 = tw[2];
    for (int j = tid; j < H; j += T) twl[j] = twH ? twH[j] : tw[3 * j];
    __syncthreads();
    auto wM = [&](int e) {
        const int q3 = e / 3, r3 = e - 3 * q3;
        const TC t = twl[q3];
        const TC m = mul(t, TC{r3 == 1 ? wm1.x : wm2.x, r3 == 1 ? wm1.y : wm2.y});
        return TC{r3 == 0 ? t.x : m.x, r3 == 0 ? t.y : m.y};
    };
    TC vel[3][PO];
    TC ra0[PJ], rb0[PJ], ra1[PJ], rb1[PJ];
    auto issue = [&](const TC *fa, TC *a, TC *b) {
        const TC *fb = fa + vstride;
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            const int q = (tid >> 6) * 32 + (lane & 31) + j * (T / 2);
            const int k = lane < 32 ? q : (q == 0 ? H / 2 : H - q);
            TC va = zero, vb = zero;
            if (q < H / 2) {
                va = fa[vp<TC>(H) * (size_t)k];
                vb = fb[vp<TC>(H) * (size_t)k];
            }
            a[j] = va;
            b[j] = vb;
        }
    };
    // u_i u_j form (DIV): one field pair (u, w), products D = u^2 - w^2, C = u w
    static_assert(!DIV || uiuj_built(H), "u_i u_j form: H = 2048");
    constexpr int NPH = DIV ? 1 : 2;
    issue(V + vrow<TC>(row_of(0), H), ra0, rb0);
#pragma unroll 1
    for (int i = 0; i < RPW; ++i) {
        // the thread index through an empty asm per row: the index arithmetic of the row
        // stays inside the loop instead of occupying registers across it
        tid = opaque((int)threadIdx.x);
        lane = tid & 63;
        const int z = row_of(i);
        const TC *row = V + vrow<TC>(z, H);
#pragma unroll
        for (int ph = 0; ph < NPH; ++ph) {
#pragma unroll
            for (int j = 0; j < PJ; ++j) {
                const int q = (tid >> 6) * 32 + (lane & 31) + j * (T / 2);
                const bool valid = q < H / 2;
                const bool hi = lane >= 32;
                const int k = !hi ? q : (q == 0 ? H / 2 : H - q);
                const TC av = ph == 0 ? ra0[j] : ra1[j], bv = ph == 0 ? rb0[j] : rb1[j];
                TC am{swap_halves(av.x), swap_halves(av.y)};
                TC bm{swap_halves(bv.x), swap_halves(bv.y)};
                if (k == H / 2) {
                    am = av;
                    bm = bv;
                }
                if (valid) {
                    TC zlo, zhi;
                    if (k == 0) {
                        zlo = TC{av.x, bv.x};
                        zhi = zero;
                    } else {
                        zlo = TC{av.x - bv.y, av.y + bv.x};
                        zhi = TC{am.x + bm.y, bm.x - am.y};
                    }
                    lds[pidx(k)] = add(zlo, zhi);
                    lds[LS + pidx(k)] = mulc(add(zlo, mulc(zhi, w3_1)), wM(k));
                    lds[2 * LS + pidx(k)] = mulc(add(zlo, mulc(zhi, w3_2)), wM(2 * k));
                }
            }
            if (!DIV && ph == 0) issue(row + 2 * vstride, ra1, rb1);
            __syncthreads();
            // in place (one barrier per stage): position p of every line holds x1 = ip_idx(p);
            // the product below is pointwise and the forward DIT takes that order back
            fft_lines_to_pos<H, LS, 3, T, true, 8, opq_row<TC>(H)>(lds, twl, 1);
#pragma unroll
            for (int j = 0; j < PO; ++j) {
                const int x1 = tid + j * T;
                if constexpr (DIV) {
#pragma unroll
                    for (int n2 = 0; n2 < 3; ++n2) {
                        const TC v = lds[n2 * LS + pidx(x1)];
                        lds[n2 * LS + pidx(x1)] = TC{v.x * v.x - v.y * v.y, v.x * v.y};
                    }
                } else if (ph == 0) {
#pragma unroll
                    for (int n2 = 0; n2 < 3; ++n2) vel[n2][j] = lds[n2 * LS + pidx(x1)];
                } else {
                    RT P[3];
#pragma unroll
                    for (int n2 = 0; n2 < 3; ++n2) {
                        const TC v = lds[n2 * LS + pidx(x1)];
                        P[n2] = vel[n2][j].x * v.x + vel[n2][j].y * v.y;
                    }
                    lds[pidx(x1)] = TC{P[0], P[1]};
                    lds[LS + pidx(x1)] = TC{P[2], RT(0)};
                }
            }
            // the next row's first pair, into the registers this row's pairs have freed: in
            // flight during the forward transform
            if (ph == NPH - 1 && i + 1 < RPW) issue(V + vrow<TC>(row_of(i + 1), H), ra0, rb0);
            __syncthreads();
        }
        const RT half = RT(0.5);
        if constexpr (DIV) {
            // x-forward of D + i C, split into D^ and C^ (as in k_row)
            fft_lines_from_pos<H, LS, 3, T, false, 8, opq_row<TC>(H)>(lds, twl, 1);
            for (int k = tid; k < H; k += T) {
                const TC X =
                    add(lds[pidx(k)], add(mul(lds[LS + pidx(k)], wM(k)), mul(lds[2 * LS + pidx(k)], wM(2 * k))));
                TC Xm = X;
                if (k != 0) {
                    const int km = H - k;
                    Xm = add(lds[pidx(km)], add(mul(mul(lds[LS + pidx(km)], wM(km)), w3_1),
                                                mul(mul(lds[2 * LS + pidx(km)], wM(2 * km)), w3_2)));
                }
                const TC zm = conj(Xm);
                const TC sm = add(X, zm), df = sub(X, zm);
                const size_t o = gpos<TC>(z, k, H);
                G[o] = TC{half * sm.x, half * sm.y};
                G[vstride + o] = TC{half * df.y, -half * df.x};
            }
            __syncthreads();  // the next row writes the lines this loop reads
            continue;
        }
        TC *A = lds;
        TC *B = lds + LS;
        fft_lines_from_pos<H, LS, 2, T, false, 8, opq_row<TC>(H)>(lds, twl, 1);
        for (int k = tid; k < H; k += T) {
            const TC za = A[pidx(k)];
            const TC zm = conj(A[pidx(k == 0 ? 0 : H - k)]);
            const TC s = add(za, zm), d = sub(za, zm);
            const TC F0{half * s.x, half * s.y};
            const TC F1{half * d.y, -half * d.x};
            const TC F2 = B[pidx(k)];
            G[gpos<TC>(z, k, H)] = add(F0, add(mul(F1, wM(k)), mul(F2, wM(2 * k))));
        }
        __syncthreads();  // the next row writes the lines this loop reads
    }
}

// =====================================================================================
// Row kernel over row groups (H = 128 .. 512): R consecutive rows per workgroup, MODE 0
// only (the step).  V is column-major, so the R rows' values of one mode kx sit side by side (16 R
// bytes in FP64); lane l of a wave takes row z0 + (l mod 32) mod R and mode slot
// (l mod 32) / R, so each load instruction reads R rows x 32 / R modes in runs of R values
// instead of one value per line.  Per row the work is k_row's, line for line (3R lines per
// transform, the same engine and twiddles), so the results are bitwise those of k_row.
// =====================================================================================
template <class TC, int H, int T, int R, int MINW, int PLAN = 0>
__global__ __launch_bounds__(T, MINW) void k_row_mr(const TC *__restrict__ V, size_t vstride, TC *__restrict__ G,
                                              const TC *__restrict__ tw, BatchArgs bt)
{
    DNS_V_PAIRS(H);
    {
        const size_t b = blockIdx.y;
        V += b * bt.s_v;
        G += b * bt.s_g;
    }
    using RT = decltype(TC::x);
    constexpr int LS = padded_len(H);
    constexpr int QW = T / (2 * R);   // modes q (pairs q, H - q) per lane sweep of the workgroup
    constexpr int PJ = (H / 2) / QW;  // pairs per lane
    constexpr int PO = R * H / T;     // outputs per thread and residue
    static_assert(T % 64 == 0 && 32 % R == 0 && (H / 2) % QW == 0 && (R * H) % T == 0, "row-group tiling");
    __shared__ TC lds[3 * R * LS];
    __shared__ TC twl[H];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int z0 = xcd_tile(blockIdx.x, gridDim.x) * R;
    const int r = (lane & 31) % R, sq = (lane & 31) / R;
    const TC *row = V + vrow<TC>(z0 + r, H);  // element (f, k) of row z0 + r at f * vstride + vp k
    const TC w3_1 = tw[2 * H], w3_2 = tw[H];
    const TC zero{RT(0), RT(0)};
    const TC wm1 = tw[1], wm2 = tw[2];
    for (int j = tid; j < H; j += T) twl[j] = tw[3 * j];
    __syncthreads();
    auto wM = [&](int e) {  // W_M^e from the table, as in k_row
        const int q3 = e / 3, r3 = e - 3 * q3;
        const TC t = twl[q3];
        const TC m = mul(t, TC{r3 == 1 ? wm1.x : wm2.x, r3 == 1 ? wm1.y : wm2.y});
        return TC{r3 == 0 ? t.x : m.x, r3 == 0 ? t.y : m.y};
    };

    TC vel[3][PO];
    TC ra0[PJ], rb0[PJ], ra1[PJ], rb1[PJ];
    auto mode_of = [&](int j, int *q) {
        *q = (tid >> 6) * (32 / R) + sq + j * QW;
        return lane < 32 ? *q : (*q == 0 ? H / 2 : H - *q);
    };
    auto issue = [&](const TC *fa, TC *a, TC *b) {
        const TC *fb = fa + vstride;
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            int q;
            const int k = mode_of(j, &q);
            a[j] = fa[vp<TC>(H) * (size_t)k];
            b[j] = fb[vp<TC>(H) * (size_t)k];
        }
    };
    // LDS line of residue n2 of row r: residue-major, all rows in one transform call
    auto line = [](int r, int n2) { return n2 * R + r; };
    issue(row, ra0, rb0);
#pragma unroll
    for (int ph = 0; ph < 2; ++ph) {
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            int q;
            const int k = mode_of(j, &q);
            const TC av = ph == 0 ? ra0[j] : ra1[j], bv = ph == 0 ? rb0[j] : rb1[j];
            TC am{swap_halves(av.x), swap_halves(av.y)};
            TC bm{swap_halves(bv.x), swap_halves(bv.y)};
            if (k == H / 2) {
                am = av;
                bm = bv;
            }
            TC zlo, zhi;
            if (k == 0) {
                zlo = TC{av.x, bv.x};
                zhi = zero;
            } else {
                zlo = TC{av.x - bv.y, av.y + bv.x};
                zhi = TC{am.x + bm.y, bm.x - am.y};
            }
            lds[line(r, 0) * LS + pidx(k)] = add(zlo, zhi);
            lds[line(r, 1) * LS + pidx(k)] = mulc(add(zlo, mulc(zhi, w3_1)), wM(k));
            lds[line(r, 2) * LS + pidx(k)] = mulc(add(zlo, mulc(zhi, w3_2)), wM(2 * k));
        }
        if (ph == 0) issue(row + 2 * vstride, ra1, rb1);
        __syncthreads();
        fft_lines_to_pos<H, LS, 3 * R, T, true, (PLAN & 2) ? 16 : 8, (PLAN & 1) != 0, ip_row<TC>(H)>(lds, twl, 1);
#pragma unroll
        for (int j = 0; j < PO; ++j) {
            const int e = tid + j * T;
            const int rr = e / H, x1 = e - rr * H;
            if (ph == 0) {
#pragma unroll
                for (int n2 = 0; n2 < 3; ++n2) vel[n2][j] = lds[line(rr, n2) * LS + pidx(x1)];
            } else {
                // the real product u d_x omega + w d_z omega, written straight into the
                // x-forward's lines A_r = P0 + i P1 (line (r, 0)) and B_r = P2 (line (r, 1)):
                // this thread alone reads and writes position x1 of row rr's lines
                RT P[3];
#pragma unroll
                for (int n2 = 0; n2 < 3; ++n2) {
                    const TC v = lds[line(rr, n2) * LS + pidx(x1)];
                    P[n2] = vel[n2][j].x * v.x + vel[n2][j].y * v.y;
                }
                lds[line(rr, 0) * LS + pidx(x1)] = TC{P[0], P[1]};
                lds[line(rr, 1) * LS + pidx(x1)] = TC{P[2], RT(0)};
            }
        }
        __syncthreads();
    }
    // x-forward of each row's product, then X[k] = F0 + W_M^k F1 + W_M^{2k} F2 for k < H
    fft_lines_from_pos<H, LS, 2 * R, T, false, (PLAN & 2) ? 16 : 8, (PLAN & 1) != 0, ip_row<TC>(H)>(lds, twl, 1);  // lines (r, 0), (r, 1) are 0 .. 2R - 1
    const RT half = RT(0.5);
    for (int e = tid; e < R * H; e += T) {
        const int rr = e / H, k = e - rr * H;
        const TC *A = lds + line(rr, 0) * LS;
        const TC za = A[pidx(k)];
        const TC zm = conj(A[pidx(k == 0 ? 0 : H - k)]);
        const TC s = add(za, zm), d = sub(za, zm);
        const TC F0{half * s.x, half * s.y};
        const TC F1{half * d.y, -half * d.x};  // d / (2 i)
        const TC F2 = lds[line(rr, 1) * LS + pidx(k)];
        G[gpos<TC>(z0 + rr, k, H)] = add(F0, add(mul(F1, wM(k)), mul(F2, wM(2 * k))));
    }
}

// =====================================================================================
// Physical plane of one or two z-inverted fields (diagnostics / output dumps).
// out_mode 0: out0 = Re, out1 = Im (if non-null); out_mode 1: out0 = |a + i b|.
// =====================================================================================
template <int H, int T>
__global__ __launch_bounds__(T) void k_row_phys(const cd *__restrict__ Va, const cd *__restrict__ Vb,
                                                const cd *__restrict__ tw, double *__restrict__ out0,
                                                double *__restrict__ out1, int out_mode)
{
    DNS_V_PAIRS(H);
    constexpr int M = 3 * H;
    constexpr int PJ = (H + T - 1) / T;
    constexpr int LS = padded_len(H);
    __shared__ cd A[LS];
    const int tid = threadIdx.x;
    const int z = blockIdx.x;
    const cd *a = Va + vrow<cd>(z, H);  // element k at vp k
    const cd *b = Vb ? Vb + vrow<cd>(z, H) : nullptr;
    for (int x2 = 0; x2 < 3; ++x2) {
        const cd w3 = tw[(2 * x2 * H) % M];
        for (int k1 = tid; k1 < H; k1 += T) {
            const cd av = a[vp<cd>(H) * (size_t)k1];
            const cd bv = b ? b[vp<cd>(H) * (size_t)k1] : mk(0.0, 0.0);
            cd zlo = (k1 == 0) ? mk(av.x, bv.x) : mk(av.x - bv.y, av.y + bv.x);
            cd zhi = mk(0.0, 0.0);
            if (k1 > 0) {
                const cd a2 = a[vp<cd>(H) * (size_t)(H - k1)];
                const cd b2 = b ? b[vp<cd>(H) * (size_t)(H - k1)] : mk(0.0, 0.0);
                zhi = mk(a2.x + b2.y, b2.x - a2.y);
            }
            A[pidx(k1)] = mulc(add(zlo, mulc(zhi, w3)), tw[opaque(x2 * k1)]);
        }
        __syncthreads();
        fft_lines<H, LS, 1, T, true, 8, false, cd>(A, tw, 3);
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            const int x1 = tid + j * T;
            if (x1 < H) {
                const cd v = A[pidx(x1)];
                const size_t o = (size_t)z * M + x2 + 3 * x1;
                if (out_mode == 1) {
                    out0[o] = sqrt(v.x * v.x + v.y * v.y);
                } else {
                    out0[o] = v.x;
                    if (out1) out1[o] = v.y;
                }
            }
        }
        __syncthreads();
    }
}

// =====================================================================================
// Reductions (diagnostics).
// =====================================================================================
template <int T>
__device__ __forceinline__ double block_sum(double v, double *red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < T / 64; ++i) s += red[i];
    __syncthreads();
    return s;
}

// E, Z partial sums per kx column (pao_km3_cuda.cu:366-378 convention)
__global__ __launch_bounds__(256) void k_energy(const cd *__restrict__ om, int N, double *__restrict__ partE,
                                                double *__restrict__ partZ)
{
    __shared__ double red[4];
    const int kx = blockIdx.x;
    const double m = (kx == 0) ? 1.0 : 2.0;
    const double ax2 = double(kx) * double(kx);
    double e = 0.0, zz = 0.0;
    for (int iz = threadIdx.x; iz < N; iz += 256) {
        const cd o = om[(size_t)kx * N + iz];
        const double gz = gamma_of(iz, N);
        const double K2 = ax2 + gz * gz;
        const double a2 = o.x * o.x + o.y * o.y;
        zz += m * a2;
        if (K2 > 0.0) e += m * a2 / K2;
    }
    const double se = block_sum<256>(e, red);
    const double sz = block_sum<256>(zz, red);
    if (threadIdx.x == 0) {
        partE[kx] = se;
        partZ[kx] = sz;
    }
}

// Palinstrophy and enstrophy partials with the ADAPT_VISC controller's weights
// (mem_cuda.cu:872-935): w = 1 on the kx = 0 and kx = H - 1 columns, 2 elsewhere; the
// (0, 0) mode is left out of the enstrophy sum.  Block b sums line b of the state:
// column kx = b of [kx][iz] (rowmajor 0, fused layout) or row iz = b of [iz][kx]
// (rowmajor 1, multi-pass layout).
__global__ __launch_bounds__(256) void k_palin(const cd *__restrict__ om, int N, int tiled,
                                               double *__restrict__ partT, double *__restrict__ partP)
{
    __shared__ double red[4];
    const int H = N / 2, b = blockIdx.x;
    // fused layout [kx][iz]: block b = column kx; multi-pass column tiles [kx/8][iz][kx%8]:
    // block b = tile of the modes 8 b .. 8 b + 7 (contiguous, N x 8 elements)
    const int len = tiled ? 8 * N : N;
    double tot = 0.0, pal = 0.0;
    for (int i = threadIdx.x; i < len; i += 256) {
        const int kx = tiled ? 8 * b + (i & 7) : b, iz = tiled ? (i >> 3) : i;
        if (kx >= H) continue;
        const cd o = om[(size_t)b * len + i];
        const double w = (kx == 0 || kx == H - 1) ? 1.0 : 2.0;
        const double gz = gamma_of(iz, N);
        const double k2 = double(kx) * double(kx) + gz * gz;
        const double p = o.x * o.x + o.y * o.y;
        if (kx != 0 || iz != 0) tot += w * p;
        pal += w * k2 * p;
    }
    const double st = block_sum<256>(tot, red);
    const double sp = block_sum<256>(pal, red);
    if (threadIdx.x == 0) {
        partT[b] = st;
        partP[b] = sp;
    }
}

__global__ __launch_bounds__(256) void k_minmax(const double *__restrict__ f, size_t n, double *__restrict__ out)
{
    __shared__ double smin[256], smax[256];
    double mn = 1.0e308, mx = -1.0e308;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double v = f[i];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            smin[threadIdx.x] = smin[threadIdx.x + o] < smin[threadIdx.x] ? smin[threadIdx.x + o] : smin[threadIdx.x];
            smax[threadIdx.x] = smax[threadIdx.x + o] > smax[threadIdx.x] ? smax[threadIdx.x + o] : smax[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = smin[0];
        out[2 * blockIdx.x + 1] = smax[0];
    }
}

// pixel statistics of computeDisplaySigma (mem_cuda.cu:691-740), FP64 quantization
__global__ __launch_bounds__(256) void k_pixels(const double *__restrict__ f, size_t n, double mn, double rng,
                                                int is_const, unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long s1[256], s2[256];
    unsigned long long a = 0, b = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        unsigned long long pix = 128;
        if (!is_const) {
            double pf = 1.0 + ((f[i] - mn) / rng) * 254.0;
            pf = pf < 1.0 ? 1.0 : (pf > 255.0 ? 255.0 : pf);
            pix = (unsigned long long)(int)pf;
        }
        a += pix;
        b += pix * pix;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = s1[0];
        out[2 * blockIdx.x + 1] = s2[0];
    }
}

// transposes between the host layout [iz][ix] and the device layout [ix][iz]
__global__ void k_transpose_cd(const cd *__restrict__ in, cd *__restrict__ out, int rows, int cols)
{
    __shared__ cd tile[16][17];
    const int bx = blockIdx.x * 16, by = blockIdx.y * 16;
    int x = bx + threadIdx.x, y = by + threadIdx.y;
    if (x < cols && y < rows) tile[threadIdx.y][threadIdx.x] = in[(size_t)y * cols + x];
    __syncthreads();
    x = by + threadIdx.x;
    y = bx + threadIdx.y;
    if (x < rows && y < cols) out[(size_t)y * rows + x] = tile[threadIdx.x][threadIdx.y];
}

// =====================================================================================
// Host-side launchers: one set of template instances per supported H.
// =====================================================================================
namespace {

template <int H>
struct Cfg;
// H: retained kx; C: columns per column-kernel workgroup; TC / TR: threads per workgroup
#define DNS_CFG(H_, C_, TC_, TR_)              \
    template <>                                \
    struct Cfg<H_> {                           \
        static constexpr int C = C_;           \
        static constexpr int TC = TC_;         \
        static constexpr int TR = TR_;         \
        static constexpr int plan = 0;         \
        static constexpr int rows = 0;         \
    };
DNS_CFG(8, 8, 64, 64)
DNS_CFG(16, 4, 64, 64)
DNS_CFG(32, 2, 64, 64)
DNS_CFG(64, 1, 64, 64)
DNS_CFG(128, 1, 64, 64)
DNS_CFG(192, 1, 192, 64)
DNS_CFG(256, 1, 128, 128)
DNS_CFG(384, 1, 192, 192)
DNS_CFG(512, 2, 512, 256)
DNS_CFG(768, 2, 768, 384)
DNS_CFG(1024, 2, 1024, 512)
DNS_CFG(1536, 1, 768, 768)
DNS_CFG(2048, 1, 1024, 1024)
#undef DNS_CFG

// Minimum waves per SIMD the compiler budgets the column kernel's VGPRs for
// (amdgpu_waves_per_eu).  Forcing two workgroups per CU (FP32 lines at H >= 1024 fit twice
// in LDS) made the kernels spill and run slower on MI355X (column 0.45 -> 0.55 ms at
// N=4096), except with the opaque FFT index: the FP32 column kernel at H = 2048 runs two
// 1024-thread workgroups per CU (8 waves per SIMD, 64 VGPRs, 72 KB of LDS each), 0.41 ->
// 0.32 ms.
template <class TC, int H, int T, int LDS_BYTES>
constexpr int col_min_waves()
{
    return ((std::is_same<TC, cf>::value || std::is_same<TC, cs>::value) && H == 2048 && T == 1024 &&
            2 * LDS_BYTES <= 160 * 1024)
               ? 8
               : 1;
}

// Launch variants (the ensemble's autotuner times them): the column tile and the
// column-kernel threads scaled by CM, the row-kernel threads by TM, PLAN (bit 0: the opaque
// FFT thread index in both kernels; bit 1: the radix-16 plan of every line transform) and
// ROWS (rows per row-kernel workgroup at H = 128 .. 512; 0: the default).  Variants without
// the radix-16 plan do the same arithmetic per element (only which thread does it changes),
// so their results are bitwise those of variant 0; the radix-16 plans round differently.
template <int H, int CM, int TM, int PLAN = 0, int ROWS = 0>
struct VarCfg {
    static constexpr int C = Cfg<H>::C * CM;
    static constexpr int TC = Cfg<H>::TC * CM;
    static constexpr int TR = Cfg<H>::TR * TM;
    static constexpr int plan = PLAN;
    static constexpr int rows = ROWS;
};
// variants are built for FP64 and H a power of two in 16 .. 512 (the ensemble sizes)
constexpr bool has_variants(int H) { return H >= 16 && H <= 512 && (H & (H - 1)) == 0; }
// the variant table: (CM, TM, PLAN, ROWS); the last kNumPlanVariants use the radix-16 plan
#define DNS_VARIANTS(X)    \
    X(0, 1, 1, 0, 0)       \
    X(1, 2, 1, 0, 0)       \
    X(2, 1, 2, 0, 0)       \
    X(3, 2, 2, 0, 0)       \
    X(4, 1, 1, 1, 0)       \
    X(5, 2, 1, 1, 0)       \
    X(6, 1, 1, 0, 4)       \
    X(7, 1, 1, 2, 0)       \
    X(8, 2, 1, 2, 0)
constexpr int kNumVariants = 9;
constexpr int kNumPlanVariants = 2;

template <class TC, int H, class Q>
hipError_t launch_cfg(const LaunchReq &r);

// rows per workgroup and threads of k_row_mr, timed on the ensemble (192 x N=256 +
// 64 x N=512): two rows per workgroup beat four and eight (fewer VGPRs and LDS, more
// workgroups per CU outweigh the longer runs of V values).  At N = 4096 in FP32, two rows
// per workgroup (127 KB of LDS, one workgroup per CU) ran 0.50 ms; the one-row kernel with
// the opaque FFT index (61 VGPRs, two 1024-thread workgroups per CU) runs 0.355 ms.
template <class TC, int H, int ROWS = 0>
struct MrCfg {
    static constexpr int R = ROWS ? ROWS : 2;
    static constexpr int T = (H == 128 ? 128 : 256) * (R / 2);
};
// which row kernel the step's row pass runs: row groups at H = 128 .. 512 (the default row
// threads), one row per workgroup otherwise
template <class TC, int H, class Q>
constexpr bool use_row_groups()
{
    return H >= 128 && H <= 512 && (H & (H - 1)) == 0 && Q::TR == Cfg<H>::TR;
}

// FP32 transforms at H = 2048 (N = 4096): the lines take half the LDS and, with the opaque
// FFT index, the kernels fit two 1024-thread workgroups per CU (row 0.50 -> 0.355 ms,
// column 0.41 -> 0.37 ms); results bitwise unchanged -- the same arithmetic per element
struct Cfg32_2048 {
    static constexpr int C = 1;
    static constexpr int TC = 1024;
    static constexpr int TR = 1024;
    static constexpr int plan = 0;
    static constexpr int rows = 0;
};

template <class TC, int H>
hipError_t launch_prec(const LaunchReq &r)
{
    if constexpr (has_variants(H) && std::is_same<TC, cd>::value) {
        switch (r.variant) {
#define DNS_VARIANT_CASE(V, CM, TM, PLAN, ROWS) \
    case V: return launch_cfg<TC, H, VarCfg<H, CM, TM, PLAN, ROWS>>(r);
            DNS_VARIANTS(DNS_VARIANT_CASE)
#undef DNS_VARIANT_CASE
            default: break;
        }
    }
    if constexpr (H == 2048 && (std::is_same<TC, cf>::value || std::is_same<TC, cs>::value))
        return launch_cfg<TC, H, Cfg32_2048>(r);
    return launch_cfg<TC, H, Cfg<H>>(r);
}

template <class TC, int H, class Q>
hipError_t launch_cfg(const LaunchReq &r)
{
    constexpr int M = 3 * H;
    // the column kernel's LDS: its 3C lines and the H-point lines' twiddle table
    constexpr int WC = col_min_waves<TC, H, Q::TC, (3 * Q::C * padded_len(H) + H) * (int)sizeof(TC)>();
    constexpr int WR = 1;
    const TC *twT = static_cast<const TC *>(r.twT);
    const unsigned nb = r.batch.count > 1 ? (unsigned)r.batch.count : 1u;
    constexpr bool S16 = std::is_same<TC, cs>::value;
    // the u_i u_j form: FP64 step kernels at H = 2048, unbatched
    constexpr bool UI = std::is_same<TC, cd>::value && uiuj_built(H) && std::is_same<Q, Cfg<H>>::value;
    if (r.form != 0) {
        if (!UI || r.form != 1 || nb != 1 || (r.what != LAUNCH_ROW && r.what != LAUNCH_COL)) return hipErrorInvalidValue;
    }
    if constexpr (S16) {
        // FP16 storage: the step's row and column passes of one solver, four V fields at most
        if (nb != 1 || !r.vexp || !r.gexp || (r.what != LAUNCH_ROW && r.what != LAUNCH_COL) ||
            (r.what == LAUNCH_COL && r.col.nfields > 4))
            return hipErrorInvalidValue;
    }
    switch (r.what) {
        case LAUNCH_COL:
            if constexpr (H == 2048 && std::is_same<TC, cd>::value && std::is_same<Q, Cfg<H>>::value) {
                // FP64 at N = 4096, the step's column pass: the persistent column kernel,
                // 256 workgroups x 8 columns
                constexpr int CPW = 8, NW = H / CPW;
                static_assert(NW % 8 == 0 && NW * CPW == H && Q::C == 1, "whole columns per workgroup");
                if (nb == 1 && !r.batch.coef && r.col.do_forward && r.col.nfields >= 1) {
                    // the u_i u_j form: the one-column kernel (the persistent one, with plane 0's
                    // share of FN held across plane 1's transform beside its prefetch, spilled and
                    // ran 0.486 against 0.480 ms)
                    if (r.form == 1)
                        hipLaunchKernelGGL((k_col<TC, H, Q::C, Q::TC, WC, Q::plan, true>), dim3(H / Q::C), dim3(Q::TC), 0,
                                           r.stream, r.om, r.fn, static_cast<const TC *>(r.G), static_cast<TC *>(r.V), twT,
                                           r.tw, r.col, r.batch, static_cast<const TC *>(r.twH), r.gexp, r.vexp);
                    else
                        hipLaunchKernelGGL((k_col_p<H, Q::TC, CPW>), dim3(NW), dim3(Q::TC), 0, r.stream, r.om, r.fn,
                                           static_cast<const TC *>(r.G), static_cast<TC *>(r.V), twT, r.tw, r.col,
                                           static_cast<const TC *>(r.twH));
                    break;
                }
            }
            hipLaunchKernelGGL((k_col<TC, H, Q::C, Q::TC, WC, Q::plan>), dim3(H / Q::C, nb), dim3(Q::TC), 0, r.stream, r.om,
                               r.fn, static_cast<const TC *>(r.G), static_cast<TC *>(r.V), twT, r.tw, r.col, r.batch,
                               static_cast<const TC *>(r.twH), r.gexp, r.vexp);
            break;
        case LAUNCH_ROW:
            // the step's row pass computes no CFL (LAUNCH_ROW_CFL does)
            if (r.cflrow) return hipErrorInvalidValue;
            // H = 128 .. 512 at the default row threads: row groups (k_row_mr); the variants
            // with doubled row threads keep one row per workgroup, so the ensemble's
            // autotuner can choose between the two
            if constexpr (use_row_groups<TC, H, Q>()) {
                constexpr int RR = MrCfg<TC, H, Q::rows>::R, TT = MrCfg<TC, H, Q::rows>::T;
                hipLaunchKernelGGL((k_row_mr<TC, H, TT, RR, 1, Q::plan>), dim3(M / RR, nb), dim3(TT), 0, r.stream,
                                   static_cast<const TC *>(r.V), r.vstride, static_cast<TC *>(r.Gw), twT, r.batch);
            } else {
                if constexpr (std::is_same<TC, cd>::value && H == 2048 && std::is_same<Q, Cfg<H>>::value) {
                    // FP64 at N = 4096: the persistent row kernel, 256 workgroups x 24 rows
                    constexpr int RPW = 24, NW = M / RPW;
                    static_assert(NW % 8 == 0 && NW * RPW == M, "whole rows per workgroup, whole XCD sets");
                    if (nb == 1) {
                        if (r.form == 1)
                            hipLaunchKernelGGL((k_row_p<TC, H, Q::TR, RPW, true>), dim3(NW), dim3(Q::TR), 0, r.stream,
                                               static_cast<const TC *>(r.V), r.vstride, static_cast<TC *>(r.Gw), twT,
                                               static_cast<const TC *>(r.twH));
                        else
                            hipLaunchKernelGGL((k_row_p<TC, H, Q::TR, RPW>), dim3(NW), dim3(Q::TR), 0, r.stream,
                                               static_cast<const TC *>(r.V), r.vstride, static_cast<TC *>(r.Gw), twT,
                                               static_cast<const TC *>(r.twH));
                        break;
                    }
                }
                hipLaunchKernelGGL((k_row<TC, H, Q::TR, 0, WR, Q::plan>), dim3(M, nb), dim3(Q::TR), 0, r.stream,
                                   static_cast<const TC *>(r.V), r.vstride, static_cast<TC *>(r.Gw), twT, r.cflrow,
                                   r.DX, r.DZ, r.batch, static_cast<const TC *>(r.twH), r.vexp, r.gexp);
            }
            break;
        case LAUNCH_ROW_CFL:
            if constexpr (!S16)
                hipLaunchKernelGGL((k_row<TC, H, Q::TR, 1, WR>), dim3(M, nb), dim3(Q::TR), 0, r.stream,
                                   static_cast<const TC *>(r.V), r.vstride, static_cast<TC *>(r.Gw), twT, r.cflrow,
                                   r.DX, r.DZ, r.batch, static_cast<const TC *>(r.twH), nullptr, nullptr);
            break;
        case LAUNCH_ROW_PHYS:
            hipLaunchKernelGGL((k_row_phys<H, Q::TR>), dim3(M), dim3(Q::TR), 0, r.stream, r.Va, r.Vb, r.tw,
                               r.out0, r.out1, r.out_mode);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// the matrix-core line transforms (cm3, cmh) are built for the N=4096 grid and the sizes
// their oracle tests use
constexpr bool mfma_built(int H) { return H == 128 || H == 256 || H == 512 || H == 1024 || H == 2048; }
// FP16 storage of V and G (cs) is built for the N=4096 grid
constexpr bool h16_built(int H) { return H == 2048; }

template <int H>
hipError_t launch_all(const LaunchReq &r)
{
    if (r.fp32) {
        if constexpr (mfma_built(H)) {
            if (r.engine == 1) return launch_prec<cm3, H>(r);
            if (r.engine == 2) return launch_prec<cmh, H>(r);
        }
        if (r.engine == 3) {
            if constexpr (h16_built(H)) return launch_prec<cs, H>(r);
            return hipErrorInvalidValue;
        }
        if (r.engine != 0) return hipErrorInvalidValue;
        if constexpr (H >= 64) return launch_prec<cf, H>(r);
        return hipErrorInvalidValue;
    }
    return launch_prec<cd, H>(r);
}

}  // namespace

int dns_kernels_variants(int H, int fp32)
{
    return (!fp32 && has_variants(H) && dns_kernels_supported(H)) ? kNumVariants : 1;
}

int dns_kernels_variant_bitwise(int v) { return (v >= 0 && v < kNumVariants - kNumPlanVariants) ? 1 : 0; }

int dns_kernels_mfma_supported(int H) { return mfma_built(H) ? 1 : 0; }

int dns_kernels_h16_supported(int H) { return h16_built(H) ? 1 : 0; }

int dns_kernels_uiuj_supported(int H) { return uiuj_built(H) ? 1 : 0; }

int dns_kernels_supported(int H)
{
    switch (H) {
        case 8: case 16: case 32: case 64: case 128: case 192: case 256: case 384:
        case 512: case 768: case 1024: case 1536: case 2048:
            return 1;
        default:
            return 0;
    }
}

int dns_kernels_col_tile(int H)
{
    switch (H) {
        case 8: return Cfg<8>::C;
        case 16: return Cfg<16>::C;
        case 32: return Cfg<32>::C;
        case 64: return Cfg<64>::C;
        case 128: return Cfg<128>::C;
        case 192: return Cfg<192>::C;
        case 256: return Cfg<256>::C;
        case 384: return Cfg<384>::C;
        case 512: return Cfg<512>::C;
        case 768: return Cfg<768>::C;
        case 1024: return Cfg<1024>::C;
        case 1536: return Cfg<1536>::C;
        case 2048: return Cfg<2048>::C;
        default: return 0;
    }
}

hipError_t dns_kernels_launch(int H, const LaunchReq &r)
{
    switch (H) {
        case 8: return launch_all<8>(r);
        case 16: return launch_all<16>(r);
        case 32: return launch_all<32>(r);
        case 64: return launch_all<64>(r);
        case 128: return launch_all<128>(r);
        case 192: return launch_all<192>(r);
        case 256: return launch_all<256>(r);
        case 384: return launch_all<384>(r);
        case 512: return launch_all<512>(r);
        case 768: return launch_all<768>(r);
        case 1024: return launch_all<1024>(r);
        case 1536: return launch_all<1536>(r);
        case 2048: return launch_all<2048>(r);
        default: return hipErrorInvalidValue;
    }
}

hipError_t dns_launch_energy(const cd *om, int N, int H, double *partE, double *partZ, hipStream_t s)
{
    hipLaunchKernelGGL(k_energy, dim3(H), dim3(256), 0, s, om, N, partE, partZ);
    return hipGetLastError();
}

hipError_t dns_launch_palin(const cd *om, int N, int tiled, double *partT, double *partP, hipStream_t s)
{
    hipLaunchKernelGGL(k_palin, dim3(tiled ? (N / 2 + 7) / 8 : N / 2), dim3(256), 0, s, om, N, tiled, partT, partP);
    return hipGetLastError();
}

hipError_t dns_launch_minmax(const double *f, size_t n, double *out, int blocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_minmax, dim3(blocks), dim3(256), 0, s, f, n, out);
    return hipGetLastError();
}

hipError_t dns_launch_pixels(const double *f, size_t n, double mn, double rng, int is_const,
                             unsigned long long *out, int blocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_pixels, dim3(blocks), dim3(256), 0, s, f, n, mn, rng, is_const, out);
    return hipGetLastError();
}

hipError_t dns_launch_transpose(const cd *in, cd *out, int rows, int cols, hipStream_t s)
{
    dim3 grid((cols + 15) / 16, (rows + 15) / 16);
    hipLaunchKernelGGL(k_transpose_cd, grid, dim3(16, 16), 0, s, in, out, rows, cols);
    return hipGetLastError();
}
