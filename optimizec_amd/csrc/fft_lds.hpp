// fft_lds.hpp -- LDS-resident Stockham FFT engine for gfx950 (wave64).
//
// A workgroup transforms NL lines of length L held in LDS (line stride LS complex
// values; FP64 `cd` or FP32 `cf`).  Each radix-R stage: every thread loads the R inputs of its butterflies
// from LDS into VGPRs (twiddled), the workgroup barriers, the thread runs the radix-R
// DFT in registers and writes the R outputs to their Stockham (self-sorting)
// positions, barrier.  The radix sequence is fixed at compile time (8, then 4, 2, 3,
// 5, 7; what remains of a length with a factor 5 or 7 goes through the prime-factor
// radices 15, 14, 35, 21, 10 first, e.g. 210 = 15 x 14), so every loop below unrolls
// and the register arrays are sized exactly.
//
// Convention: forward = e^{-2 pi i k n / L}, unnormalized; INV selects e^{+...}.
// Twiddles come from one table tw[t] = e^{-2 pi i t / M} (t in [0, M)) shared by all
// line lengths L that divide M: W_L^s = tw[s * (M / L)].
#pragma once

#include <hip/hip_runtime.h>

#include "dns_types.h"

namespace dnsfft {

__device__ __forceinline__ cd mk(double a, double b) { return cd{a, b}; }
__device__ __forceinline__ cf mkf(float a, float b) { return cf{a, b}; }
template <class C>
__device__ __forceinline__ C add(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <class C>
__device__ __forceinline__ C sub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <class C>
__device__ __forceinline__ C mul(C a, C b)
{
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class C>
__device__ __forceinline__ C mulc(C a, C b)  // a * conj(b)
{
    return C{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <class C, class S>
__device__ __forceinline__ C scl(C a, S s)
{
    using RT = decltype(a.x);
    return C{a.x * RT(s), a.y * RT(s)};
}
template <class C>
__device__ __forceinline__ C conj(C a) { return C{a.x, -a.y}; }
// multiply by +i (S>0) or -i (S<0)
template <int S, class C>
__device__ __forceinline__ C muli(C a)
{
    return S > 0 ? C{-a.y, a.x} : C{a.y, -a.x};
}
__device__ __forceinline__ cf to_f(cd a) { return cf{float(a.x), float(a.y)}; }
__device__ __forceinline__ cd to_d(cf a) { return cd{double(a.x), double(a.y)}; }
__device__ __forceinline__ cd to_d(cm3 a) { return cd{double(a.x), double(a.y)}; }
__device__ __forceinline__ cd to_d(cmh a) { return cd{double(a.x), double(a.y)}; }
__device__ __forceinline__ cd to_d(cs a) { return cd{double(a.x), double(a.y)}; }
__device__ __forceinline__ cd to_d(cd a) { return a; }
template <class C>
__device__ __forceinline__ C from_d(cd a);
template <>
__device__ __forceinline__ cd from_d<cd>(cd a) { return a; }
template <>
__device__ __forceinline__ cf from_d<cf>(cd a) { return to_f(a); }
template <>
__device__ __forceinline__ cm3 from_d<cm3>(cd a) { return cm3{float(a.x), float(a.y)}; }
template <>
__device__ __forceinline__ cmh from_d<cmh>(cd a) { return cmh{float(a.x), float(a.y)}; }
template <>
__device__ __forceinline__ cs from_d<cs>(cd a) { return cs{float(a.x), float(a.y)}; }

// 0: vector radix stages; 1: three-BF16-product DFT-matrix stages (cm3); 2: FP16 (cmh)
template <class C>
struct mfma_kind {
    static constexpr int value = 0;
};
template <>
struct mfma_kind<cm3> {
    static constexpr int value = 1;
};
template <>
struct mfma_kind<cmh> {
    static constexpr int value = 2;
};

// Returns x through an empty asm statement: table loads indexed by the result cannot
// be hoisted out of loops by the compiler (twiddles are loop-invariant in many callers,
// and keeping all of them live costs more VGPRs than reloading them from L1/L2).
__device__ __forceinline__ int opaque(int x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// Value of lane l ^ 32 of the wavefront (the other half).  Doubles go through
// v_permlane32_swap, a VALU lane swap, instead of the two LDS permutes of a 64-bit
// __shfl_xor (FP64 row kernels at N = 4096: 0.713 -> 0.706 ms; the ensemble step -1%); the
// swap of a register with itself leaves lanes 32-63 of the first result and lanes 0-31 of the
// second holding the other half's values.  Floats keep the one LDS permute (the FP32 row
// kernel timed 2% slower with the swap).  Blocks of whole wavefronts (lane = threadIdx.x mod 64).
__device__ __forceinline__ unsigned swap_halves_u32(unsigned x, bool upper)
{
    const auto s = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return upper ? s[0] : s[1];
}
__device__ __forceinline__ double swap_halves(double x)
{
    const bool upper = (threadIdx.x & 32) != 0;
    const unsigned hi = swap_halves_u32((unsigned)__double2hiint(x), upper);
    const unsigned lo = swap_halves_u32((unsigned)__double2loint(x), upper);
    return __hiloint2double((int)hi, (int)lo);
}
__device__ __forceinline__ float swap_halves(float x) { return __shfl_xor(x, 32, 64); }

// LDS position of element i of a line: one 16-byte pad after every 8 elements, so
// the stride-8 and stride-64 accesses of the radix-8 stages spread over all banks.
__host__ __device__ constexpr int pidx(int i) { return i + (i >> 3); }
// padded storage (complex elements) of a line of L elements
__host__ __device__ constexpr int padded_len(int L) { return L + (L + 7) / 8; }

// ---- compile-time radix plan -------------------------------------------------
// Lengths with a factor 5 or 7 take the composite radices first (fewer LDS passes and
// barriers: 210 runs as 15 x 14 instead of 2 x 3 x 5 x 7; timed against 14 x 15 and
// 10 x 21 on the 210-point lines of the 44100-point transform).
__host__ __device__ constexpr int pick_radix57(int m)
{
    return (m % 15 == 0) ? 15 : (m % 14 == 0) ? 14 : (m % 35 == 0) ? 35 : (m % 21 == 0) ? 21 : (m % 10 == 0) ? 10 : 0;
}
// Radix plans.  maxr = 8 (the default): radix 8 while 8 divides, then 4, the composites,
// 2, 3, 5, 7.  maxr = 16: the power-of-two part 2^p (p >= 3) runs in the fewest stages of
// radix <= 16, ceil(p / 4), with the radices as even as possible, largest first
// (2048 = 16 x 16 x 8, 1024 = 16 x 8 x 8, 512 = 8 x 8 x 8, 128 = 16 x 8): fewer LDS passes
// and barriers for more registers per butterfly, which pays only in some kernels (chosen
// by timing where the engine is called).
__host__ __device__ constexpr int pow2_radix(int m, int maxr)
{
    if (maxr < 16) return 8;
    int p = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++p;
    }
    const int st = (p + 3) / 4;
    return 1 << ((p + st - 1) / st);
}
__host__ __device__ constexpr int pick_radix(int m, int maxr = 8)
{
    return (m % 8 == 0) ? pow2_radix(m, maxr) : (m % 4 == 0) ? 4 : (pick_radix57(m) != 0) ? pick_radix57(m)
         : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : (m % 5 == 0) ? 5 : (m % 7 == 0) ? 7 : m;
}
__host__ __device__ constexpr int stage_radix(int L, int s, int maxr = 8)
{
    int m = L;
    for (int i = 0; i < s; ++i) m /= pick_radix(m, maxr);
    return m == 1 ? 1 : pick_radix(m, maxr);
}
__host__ __device__ constexpr int stage_span(int L, int s, int maxr = 8)  // product of earlier radices
{
    int ns = 1, m = L;
    for (int i = 0; i < s; ++i) {
        const int r = pick_radix(m, maxr);
        ns *= r;
        m /= r;
    }
    return ns;
}
__host__ __device__ constexpr int num_stages(int L, int maxr = 8)
{
    int n = 0, m = L;
    while (m > 1) {
        m /= pick_radix(m, maxr);
        ++n;
    }
    return n;
}
__host__ __device__ constexpr bool smooth7(int L)
{
    int m = L;
    while (m > 1) {
        const int r = pick_radix(m);
        if (r > 8 && r != 10 && r != 14 && r != 15 && r != 21 && r != 35) return false;
        m /= r;
    }
    return true;
}

// ---- small DFTs in registers ----------------------------------------------------
// Constants are the exact values cos/sin(2 pi q / R) to double precision.
template <int R1, int R2, bool INV, class C>
__device__ __forceinline__ void dft_pfa(C *v);

template <int R, bool INV, class C>
__device__ __forceinline__ void dft(C *v)
{
    using RT = decltype(v[0].x);
    constexpr int S = INV ? 1 : -1;  // sign of the exponent
    if constexpr (R == 2) {
        const C a = v[0], b = v[1];
        v[0] = add(a, b);
        v[1] = sub(a, b);
    } else if constexpr (R == 4) {
        const C s02 = add(v[0], v[2]), d02 = sub(v[0], v[2]);
        const C s13 = add(v[1], v[3]), d13 = muli<S>(sub(v[1], v[3]));
        v[0] = add(s02, s13);
        v[1] = add(d02, d13);
        v[2] = sub(s02, s13);
        v[3] = sub(d02, d13);
    } else if constexpr (R == 8) {
        C e[4] = {v[0], v[2], v[4], v[6]};
        C o[4] = {v[1], v[3], v[5], v[7]};
        dft<4, INV, C>(e);
        dft<4, INV, C>(o);
        constexpr double h = 0.70710678118654752440084436210484903928;
        const C t1 = mul(o[1], C{RT(h), RT(S * h)});
        const C t2 = muli<S>(o[2]);
        const C t3 = mul(o[3], C{RT(-h), RT(S * h)});
        v[0] = add(e[0], o[0]);
        v[4] = sub(e[0], o[0]);
        v[1] = add(e[1], t1);
        v[5] = sub(e[1], t1);
        v[2] = add(e[2], t2);
        v[6] = sub(e[2], t2);
        v[3] = add(e[3], t3);
        v[7] = sub(e[3], t3);
    } else if constexpr (R == 3) {
        constexpr double c1 = -0.5;
        constexpr double s1 = 0.86602540378443864676372317075293618347;
        const C sp = add(v[1], v[2]), dm = sub(v[1], v[2]);
        const C t = add(v[0], scl(sp, c1));
        const C u = muli<S>(scl(dm, s1));
        v[0] = add(v[0], sp);
        v[1] = add(t, u);
        v[2] = sub(t, u);
    } else if constexpr (R == 5) {
        constexpr double c1 = 0.30901699437494742410229341718281905886;
        constexpr double c2 = -0.80901699437494742410229341718281905886;
        constexpr double s1 = 0.95105651629515357211643933337938214340;
        constexpr double s2 = 0.58778525229247312916870595463907276860;
        const C sp1 = add(v[1], v[4]), dm1 = sub(v[1], v[4]);
        const C sp2 = add(v[2], v[3]), dm2 = sub(v[2], v[3]);
        const C t1 = add(v[0], add(scl(sp1, c1), scl(sp2, c2)));
        const C t2 = add(v[0], add(scl(sp1, c2), scl(sp2, c1)));
        const C u1 = muli<S>(add(scl(dm1, s1), scl(dm2, s2)));
        const C u2 = muli<S>(sub(scl(dm1, s2), scl(dm2, s1)));
        v[0] = add(v[0], add(sp1, sp2));
        v[1] = add(t1, u1);
        v[4] = sub(t1, u1);
        v[2] = add(t2, u2);
        v[3] = sub(t2, u2);
    } else if constexpr (R == 7) {
        constexpr double c1 = 0.62348980185873353052500488400423981063;
        constexpr double c2 = -0.22252093395631440428890256449679475947;
        constexpr double c3 = -0.90096886790241912623610231950744505116;
        constexpr double s1 = 0.78183148246802980870844452667405775023;
        constexpr double s2 = 0.97492791218182360701813168299393121723;
        constexpr double s3 = 0.43388373911755812047576833284835875461;
        const C sp1 = add(v[1], v[6]), dm1 = sub(v[1], v[6]);
        const C sp2 = add(v[2], v[5]), dm2 = sub(v[2], v[5]);
        const C sp3 = add(v[3], v[4]), dm3 = sub(v[3], v[4]);
        // q = 1, 2, 3: cos(2 pi r q / 7) and sin(2 pi r q / 7) for r = 1, 2, 3
        const C t1 = add(v[0], add(add(scl(sp1, c1), scl(sp2, c2)), scl(sp3, c3)));
        const C t2 = add(v[0], add(add(scl(sp1, c2), scl(sp2, c3)), scl(sp3, c1)));
        const C t3 = add(v[0], add(add(scl(sp1, c3), scl(sp2, c1)), scl(sp3, c2)));
        const C u1 = muli<S>(add(add(scl(dm1, s1), scl(dm2, s2)), scl(dm3, s3)));
        const C u2 = muli<S>(sub(sub(scl(dm1, s2), scl(dm2, s3)), scl(dm3, s1)));
        const C u3 = muli<S>(add(sub(scl(dm1, s3), scl(dm2, s1)), scl(dm3, s2)));
        v[0] = add(v[0], add(add(sp1, sp2), sp3));
        v[1] = add(t1, u1);
        v[6] = sub(t1, u1);
        v[2] = add(t2, u2);
        v[5] = sub(t2, u2);
        v[3] = add(t3, u3);
        v[4] = sub(t3, u3);
    } else if constexpr (R == 16) {
        // 4 x 4: DFT-4 over n1 of x[4 n1 + n2] for each n2, times W_16^(n2 k1), then DFT-4
        // over n2 for each k1 -> X[k1 + 4 k2]
        constexpr double c1 = 0.92387953251128675612818318939678828682;  // cos(pi / 8)
        constexpr double s1 = 0.38268343236508977172845998403039886676;  // sin(pi / 8)
        constexpr double h = 0.70710678118654752440084436210484903928;
        C y[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            C t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
            dft<4, INV, C>(t);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = t[k1];
        }
        // W_16^e = cos(2 pi e / 16) + i S sin(2 pi e / 16) for e = n2 k1 in {1, 2, 3, 4, 6, 9}
        y[1][1] = mul(y[1][1], C{RT(c1), RT(S * s1)});
        y[1][2] = mul(y[1][2], C{RT(h), RT(S * h)});
        y[1][3] = mul(y[1][3], C{RT(s1), RT(S * c1)});
        y[2][1] = mul(y[2][1], C{RT(h), RT(S * h)});
        y[2][2] = muli<S>(y[2][2]);
        y[2][3] = mul(y[2][3], C{RT(-h), RT(S * h)});
        y[3][1] = mul(y[3][1], C{RT(s1), RT(S * c1)});
        y[3][2] = mul(y[3][2], C{RT(-h), RT(S * h)});
        y[3][3] = mul(y[3][3], C{RT(-c1), RT(-S * s1)});
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            C t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
            dft<4, INV, C>(t);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
        }
    } else if constexpr (R == 10) {
        dft_pfa<2, 5, INV, C>(v);
    } else if constexpr (R == 14) {
        dft_pfa<2, 7, INV, C>(v);
    } else if constexpr (R == 15) {
        dft_pfa<3, 5, INV, C>(v);
    } else if constexpr (R == 21) {
        dft_pfa<3, 7, INV, C>(v);
    } else if constexpr (R == 35) {
        dft_pfa<5, 7, INV, C>(v);
    } else {
        static_assert(R == 2, "radix without a register DFT");
    }
}

constexpr int inv_mod(int a, int m)
{
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return 0;
}

// Prime-factor (Good-Thomas) DFT of length R1 R2, gcd(R1, R2) = 1, without twiddles:
// input n = (n1 R2 + n2 R1) mod R, output k = (k1 R2 (R2^-1 mod R1) + k2 R1 (R1^-1 mod R2))
// mod R, and X[k] = sum_{n1} W_R1^{n1 k1} sum_{n2} W_R2^{n2 k2} x[n].  Every index is a
// compile-time constant, so the permutations cost nothing.
template <int R1, int R2, bool INV, class C>
__device__ __forceinline__ void dft_pfa(C *v)
{
    constexpr int R = R1 * R2;
    constexpr int E1 = R2 * inv_mod(R2 % R1, R1), E2 = R1 * inv_mod(R1 % R2, R2);
    C t[R1][R2];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1)
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) t[n1][n2] = v[(n1 * R2 + n2 * R1) % R];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
        C c[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) c[n1] = t[n1][n2];
        dft<R1, INV, C>(c);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) t[k1][n2] = c[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) dft<R2, INV, C>(t[k1]);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) v[(k1 * E1 + k2 * E2) % R] = t[k1][k2];
}

// ---- one Stockham stage over NL lines in LDS -----------------------------------
template <int L, int LS, int NL, int T, int R, int NS, bool INV, bool OPQ, class C>
__device__ __forceinline__ void stage(C *buf, const C *__restrict__ tw, int twmul)
{
    constexpr int LR = L / R;
    constexpr int NB = NL * LR;
    constexpr int PER = (NB + T - 1) / T;
    constexpr int TWS = L / (NS * R);
    C v[PER][R];
    // OPQ: the thread index through an empty asm, so the stage's index arithmetic is not
    // hoisted out of a caller's loop into registers (fewer VGPRs; faster in some kernels)
    const int tid = OPQ ? opaque((int)threadIdx.x) : (int)threadIdx.x;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int b = tid + q * T;
        if ((NB % T) == 0 || b < NB) {
            const int line = b / LR;
            const int j = b - line * LR;
            const int k = j % NS;
            const C *src = buf + line * LS;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = src[pidx(j + r * LR)];
            if (NS > 1) {
                // w^r for r = 1..R-1 from one table load: w^2 = w w, w^(2m) = (w^m)^2,
                // w^(2m+1) = w^(2m) w (at most 3 roundings deep for r < 8); radices above
                // 8 load w^8 too and take w^r = w^8 w^(r-8) for r >= 8 (4 roundings deep)
                const C w1 = tw[opaque(k * TWS * twmul)];
                C wp[R];
                wp[1] = w1;
#pragma unroll
                for (int r = 2; r < R && r < 8; ++r) wp[r] = (r % 2 == 0) ? mul(wp[r / 2], wp[r / 2]) : mul(wp[r - 1], w1);
                if constexpr (R > 8) {
                    const C w8 = tw[opaque(8 * k * TWS * twmul)];
                    wp[8] = w8;
#pragma unroll
                    for (int r = 9; r < R; ++r) wp[r] = mul(w8, wp[r - 8]);
                }
#pragma unroll
                for (int r = 1; r < R; ++r) v[q][r] = INV ? mulc(v[q][r], wp[r]) : mul(v[q][r], wp[r]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int b = tid + q * T;
        if ((NB % T) == 0 || b < NB) {
            const int line = b / LR;
            const int j = b - line * LR;
            const int k = j % NS;
            dft<R, INV, C>(v[q]);
            C *dst = buf + line * LS;
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[pidx(base + r * NS)] = v[q][r];
        }
    }
    __syncthreads();
}

// ---- one Stockham stage as a DFT-matrix product on the matrix cores ----------------
// The radix-R DFTs (R <= 16) of 16 butterflies at a time are one product Y = F X on
// v_mfma_f32_16x16x32_{bf16,f16} (one wavefront; FP32 accumulation), in real form:
//   columns n (16)  : butterflies
//   rows m (16)     : output y = 8 t + m / 2 of row tile t (R > 8: t = 0, 1), Re (m even)
//                     or Im (m odd)
//   k (32)          : input x = 4 (k / 8) + (k mod 8) / 2, Re (k even) or Im (k odd)
// A[m][k] is Re/Im of W_R^{+-x y}; lane l supplies A row l % 16 and B column l % 16 at
// k = 8 (l / 16) .. + 7, and receives D rows 4 (l / 16) .. + 3 of column l % 16, i.e. the
// complex outputs y = 8 t + 2 (l / 16) and y + 1 of its butterfly (operand layout as
// documented for the 16 x 16 x 32 instruction; the transforms are checked against the
// FP64 oracle).  The stage's twiddles multiply the inputs in FP32 before the product.
//   cm3: A = A_hi + A_lo and X = X_hi + X_lo in BF16 (8-bit mantissas, 16 together);
//        Y = A_hi X_hi + A_hi X_lo + A_lo X_hi (the A_lo X_lo term, ~2^-18 relative, is
//        dropped).  R <= 8 packs X_hi (k < 16) and X_lo (k >= 16) into one operand and
//        takes [A_hi | A_hi] and [A_lo | 0] against it: two products instead of three.
//   cmh: each butterfly's inputs are scaled by 2^-e (e: exponent of their largest
//        component, shared by the four lanes of the column) so they lie in (-1, 1), then
//        rounded to FP16 (11-bit significands); one product; the outputs are scaled by 2^e.
typedef __bf16 mf_bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 mf_f16x8 __attribute__((ext_vector_type(8)));
typedef float mf_f32x4 __attribute__((ext_vector_type(4)));

template <int L, int LS, int NL, int T, int R, int NS, bool INV, bool OPQ, class C>
__device__ __forceinline__ void stage_mfma(C *buf, const C *__restrict__ tw, int twmul)
{
    constexpr int KIND = mfma_kind<C>::value;
    static_assert(KIND == 1 || KIND == 2, "matrix-core stage for cm3 / cmh lines");
    static_assert(R >= 2 && R <= 16 && T % 64 == 0, "radix <= 16, whole wavefronts");
    constexpr int LR = L / R;
    constexpr int NB = NL * LR;              // butterflies
    constexpr int NT = (NB + 15) / 16;       // tiles of 16 butterflies
    constexpr int NW = T / 64;
    constexpr int PT = (NT + NW - 1) / NW;   // tiles per wavefront
    constexpr int TWS = L / (NS * R);
    constexpr bool TWO = R > 8;              // two row tiles
    const int tid = OPQ ? opaque((int)threadIdx.x) : (int)threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int n = lane & 15, q = lane >> 4;
    const int x0 = TWO ? 4 * q : 4 * (q & 1);  // first of this lane's 4 inputs

    // DFT-matrix operands (per lane, per stage): W_R^e = W_L^{e L / R} from the table, one
    // load per distinct (x, y) of the lane (4 per row tile; R = 16: row tile 1 from row tile
    // 0, W_16^{x (y + 8)} = (-1)^x W_16^{x y})
    mf_bf16x8 ah[2], al[2];
    mf_f16x8 af[2];
    C wxy[2][4];
#pragma unroll
    for (int t = 0; t < (TWO ? 2 : 1); ++t) {
        const int y = 8 * t + (n >> 1);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int x = x0 + j;
            C w{0.0f, 0.0f};
            if (R == 16 && t == 1) {
                w = (x & 1) ? C{-wxy[0][j].x, -wxy[0][j].y} : wxy[0][j];
            } else if (x < R && y < R) {
                w = tw[opaque(((x * y) % R) * (L / R) * twmul)];
            }
            wxy[t][j] = w;
        }
    }
#pragma unroll
    for (int t = 0; t < (TWO ? 2 : 1); ++t) {
        const int y = 8 * t + (n >> 1), c = n & 1;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int x = x0 + (i >> 1), d = i & 1;
            float coef = 0.0f;
            if (x < R && y < R) {
                const C w = wxy[t][i >> 1];
                const float wi = INV ? -w.y : w.y;
                coef = (c == 0) ? (d == 0 ? w.x : -wi) : (d == 0 ? wi : w.x);
            }
            if constexpr (KIND == 1) {
                const __bf16 h = (__bf16)coef;
                ah[t][i] = h;
                // R <= 8: lanes q >= 2 meet X_lo, which only A_hi multiplies
                al[t][i] = (!TWO && q >= 2) ? (__bf16)0.0f : (__bf16)(coef - (float)h);
            } else {
                af[t][i] = (_Float16)coef;
            }
        }
    }

    // inputs of this lane's butterflies, twiddled, as matrix operands
    mf_bf16x8 bh[PT], bl[PT];
    mf_f16x8 bf[PT];
    int ex[PT];
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        const int tile = wv + p * NW;
        const int b = tile * 16 + n;
        float v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = 0.0f;
        if (tile < NT && b < NB) {
            const int line = b / LR;
            const int j = b - line * LR;
            const int k = j % NS;
            const C *src = buf + line * LS;
#pragma unroll
            for (int i2 = 0; i2 < 4; ++i2) {
                const int r = x0 + i2;
                if (r < R) {
                    C xv = src[pidx(j + r * LR)];
                    if (NS > 1 && r > 0) {
                        const C w = tw[opaque(r * k * TWS * twmul)];
                        xv = INV ? mulc(xv, w) : mul(xv, w);
                    }
                    v[2 * i2] = xv.x;
                    v[2 * i2 + 1] = xv.y;
                }
            }
        }
        if constexpr (KIND == 1) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const __bf16 h = (__bf16)v[i];
                const __bf16 l = (__bf16)(v[i] - (float)h);
                if constexpr (TWO) {
                    bh[p][i] = h;
                    bl[p][i] = l;
                } else {
                    bh[p][i] = q < 2 ? h : l;  // [X_hi | X_lo] in one operand
                }
            }
        } else {
            float m = 0.0f;
#pragma unroll
            for (int i = 0; i < 8; ++i) m = fmaxf(m, fabsf(v[i]));
            m = fmaxf(m, __shfl_xor(m, 16, 64));
            m = fmaxf(m, __shfl_xor(m, 32, 64));
            int e = 0;
            (void)frexpf(m, &e);  // m < 2^e (e = 0 for m = 0)
            ex[p] = e;
#pragma unroll
            for (int i = 0; i < 8; ++i) bf[p][i] = (!TWO && q >= 2) ? (_Float16)0.0f : (_Float16)ldexpf(v[i], -e);
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        const int tile = wv + p * NW;
        if (tile < NT) {  // uniform over the wavefront
            mf_f32x4 acc[2];
#pragma unroll
            for (int t = 0; t < (TWO ? 2 : 1); ++t) {
                acc[t] = mf_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                if constexpr (KIND == 1) {
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[t], bh[p], acc[t], 0, 0, 0);
                    if constexpr (TWO) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[t], bl[p], acc[t], 0, 0, 0);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[t], bh[p], acc[t], 0, 0, 0);
                } else {
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(af[t], bf[p], acc[t], 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < 4; ++i) acc[t][i] = ldexpf(acc[t][i], ex[p]);
                }
            }
            const int b = tile * 16 + n;
            if (b < NB) {
                const int line = b / LR;
                const int j = b - line * LR;
                const int k = j % NS;
                C *dst = buf + line * LS;
                const int base = (j / NS) * NS * R + k;
#pragma unroll
                for (int t = 0; t < (TWO ? 2 : 1); ++t) {
                    const int y0 = 8 * t + 2 * q;
                    if (y0 < R) dst[pidx(base + y0 * NS)] = C{acc[t][0], acc[t][1]};
                    if (y0 + 1 < R) dst[pidx(base + (y0 + 1) * NS)] = C{acc[t][2], acc[t][3]};
                }
            }
        }
    }
    __syncthreads();
}

// ---- in-place stages: decimation in frequency (DIF) and in time (DIT) ---------------
// Stage s of a plan (radices R_0 .. R_{S-1} from pick_radix) has the stride
// ST_s = L / (R_0 ... R_s).  Butterfly (blk, j), j < ST_s, owns the positions
// blk R_s ST_s + j + r ST_s (r < R_s) of its line and writes its outputs back to them, so a
// stage needs one barrier (after it) where a Stockham stage needs two.
//   DIF, stages s = 0 .. S-1: DFT over r, then output r times W_{R_s ST_s}^{j r}.  Natural
//        input; index k of the transform ends at position ip_pos(k).
//   DIT, stages s = S-1 .. 0: input r times W_{R_s ST_s}^{j r}, then DFT over r.  Index k of
//        the input starts at position ip_pos(k); natural output.  (DIT is DIF transposed,
//        and the DFT matrix is symmetric.)
// A DIF inverse followed by a pointwise operation and a DIT forward therefore never
// permutes anything: position p holds index ip_idx(p) in both.
__host__ __device__ constexpr int ip_stride(int L, int s, int maxr = 8)
{
    return L / (stage_span(L, s, maxr) * stage_radix(L, s, maxr));
}
// k = d_0 + R_0 (d_1 + R_1 (d_2 + ...))  ->  position sum_s d_s ST_s
__host__ __device__ constexpr int ip_pos(int L, int k, int maxr = 8)
{
    int pos = 0, m = L;
    while (m > 1) {
        const int r = pick_radix(m, maxr);
        m /= r;
        pos += (k % r) * m;
        k /= r;
    }
    return pos;
}
// the index held at position p (inverse of ip_pos)
__host__ __device__ constexpr int ip_idx(int L, int p, int maxr = 8)
{
    int k = 0, m = L, span = 1;
    while (m > 1) {
        const int r = pick_radix(m, maxr);
        m /= r;
        k += (p / m) * span;
        p %= m;
        span *= r;
    }
    return k;
}

// One in-place butterfly's arithmetic on its R values; w1 = W_BL^j and, for R > 8,
// w8 = W_BL^{8 j} (forward sign; the inverse takes conjugates).  Powers as in stage().
template <int R, bool INV, class C>
__device__ __forceinline__ void twiddle_ip(C *v, C w1, C w8)  // v[r] *= w^r (conj for INV)
{
    C wp[R];
    wp[1] = w1;
#pragma unroll
    for (int r = 2; r < R && r < 8; ++r) wp[r] = (r % 2 == 0) ? mul(wp[r / 2], wp[r / 2]) : mul(wp[r - 1], w1);
    if constexpr (R > 8) {
        wp[8] = w8;
#pragma unroll
        for (int r = 9; r < R; ++r) wp[r] = mul(w8, wp[r - 8]);
    }
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = INV ? mulc(v[r], wp[r]) : mul(v[r], wp[r]);
}
template <int R, bool INV, bool DIT, bool TW, class C>
__device__ __forceinline__ void bfly_ip(C *v, C w1, C w8)
{
    if constexpr (DIT && TW) twiddle_ip<R, INV, C>(v, w1, w8);
    dft<R, INV, C>(v);
    if constexpr (!DIT && TW) twiddle_ip<R, INV, C>(v, w1, w8);
}
// x returned through an empty asm that also takes dep: whatever is indexed by the result
// waits for dep to be computed (keeps a DIF butterfly's twiddles from being loaded and
// powered up while its DFT still holds all its registers)
template <class C>
__device__ __forceinline__ int after(int x, C dep)
{
    asm volatile("" : "+v"(x) : "v"(dep.x));
    return x;
}

template <int L, int LS, int NL, int T, int R, int ST, bool INV, bool DIT, bool OPQ, class C>
__device__ __forceinline__ void stage_ip(C *buf, const C *__restrict__ tw, int twmul)
{
    constexpr int LR = L / R;  // butterflies per line
    constexpr int NB = NL * LR;
    constexpr int PER = (NB + T - 1) / T;
    constexpr int BL = R * ST;
    constexpr int TWS = L / BL;  // W_BL^s = W_L^{s TWS}
    const int tid = OPQ ? opaque((int)threadIdx.x) : (int)threadIdx.x;
    // one butterfly after the other (a thread's later loads cannot pass its earlier stores
    // to the same buffer; other waves fill the gaps)
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int b = tid + q * T;
        if ((NB % T) == 0 || b < NB) {
            const int line = b / LR;
            const int jj = b - line * LR;
            const int blk = jj / ST, j = jj - blk * ST;
            C v[R];
            {
                const C *src = buf + line * LS;
                const int base = blk * BL + j;
#pragma unroll
                for (int r = 0; r < R; ++r) v[r] = src[pidx(base + r * ST)];
            }
            C w1{}, w8{};
            if constexpr (DIT || ST == 1) {
                if constexpr (ST > 1) {
                    w1 = tw[opaque(j * TWS * twmul)];
                    if constexpr (R > 8) w8 = tw[opaque(8 * j * TWS * twmul)];
                }
                bfly_ip<R, INV, DIT, (ST > 1), C>(v, w1, w8);
            } else {  // DIF: the twiddles after the DFT, see after()
                dft<R, INV, C>(v);
                const int e = after(j * TWS * twmul, v[0]);
                w1 = tw[e];
                if constexpr (R > 8) w8 = tw[opaque(8 * e)];
                twiddle_ip<R, INV, C>(v, w1, w8);
            }
            C *dst = buf + line * LS;
            const int base = blk * BL + j;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[pidx(base + r * ST)] = v[r];
        }
    }
    __syncthreads();
}

template <int L, int LS, int NL, int T, bool INV, bool DIT, int S, int MAXR, bool OPQ, class C>
__device__ __forceinline__ void stages_ip_from(C *buf, const C *__restrict__ tw, int twmul)
{
    constexpr int NSTG = num_stages(L, MAXR);
    if constexpr (S < NSTG) {
        constexpr int s = DIT ? NSTG - 1 - S : S;
        stage_ip<L, LS, NL, T, stage_radix(L, s, MAXR), ip_stride(L, s, MAXR), INV, DIT, OPQ, C>(buf, tw, twmul);
        stages_ip_from<L, LS, NL, T, INV, DIT, S + 1, MAXR, OPQ, C>(buf, tw, twmul);
    }
}

// In-place counterpart of fft_lines (vector stages only): DIT = false transforms natural
// input into the ip_pos order, DIT = true transforms ip_pos-ordered input into natural
// order.  Same conventions as fft_lines otherwise (callers barrier before; the engine ends
// with a barrier).
template <int L, int LS, int NL, int T, bool INV, bool DIT, int MAXR = 8, bool OPQ = false, class C>
__device__ __forceinline__ void fft_lines_ip(C *buf, const C *__restrict__ tw, int twmul)
{
    static_assert(mfma_kind<C>::value == 0, "in-place stages: vector radices only");
    static_assert(smooth7(L), "line length must factor into 2, 3, 5, 7");
    static_assert(LS >= padded_len(L), "line stride must hold the padded line");
    static_assert(MAXR == 8 || MAXR == 16, "radix plans: 8 or 16");
    stages_ip_from<L, LS, NL, T, INV, DIT, 0, MAXR, OPQ, C>(buf, tw, twmul);
}

// Line pairs that only need "position p of the inverse's output is position p of the
// forward's input" (pointwise work in between): in-place stages (DIF inverse / DIT forward)
// for vector lines where the caller asks for them (IP), Stockham stages otherwise (natural
// order on both sides; always for matrix-core lines).  lines_in_place<C, IP>() tells the
// caller which order the forward's input is in.
template <class C, bool IP = true>
__host__ __device__ constexpr bool lines_in_place()
{
    return IP && mfma_kind<C>::value == 0;
}
template <int L, int LS, int NL, int T, bool INV, int S, int MAXR, bool OPQ, class C>
__device__ __forceinline__ void stages_from(C *buf, const C *__restrict__ tw, int twmul);
template <int L, int LS, int NL, int T, bool INV, int MAXR = 8, bool OPQ = false, class C>
__device__ __forceinline__ void fft_lines(C *buf, const C *__restrict__ tw, int twmul);
// natural input -> position order (in place) / natural order (Stockham)
template <int L, int LS, int NL, int T, bool INV, int MAXR = 8, bool OPQ = false, bool IP = true, class C>
__device__ __forceinline__ void fft_lines_to_pos(C *buf, const C *__restrict__ tw, int twmul)
{
    if constexpr (lines_in_place<C, IP>())
        fft_lines_ip<L, LS, NL, T, INV, false, MAXR, OPQ>(buf, tw, twmul);
    else
        fft_lines<L, LS, NL, T, INV, MAXR, OPQ>(buf, tw, twmul);
}
// position-order input (in place) / natural input (Stockham) -> natural order
template <int L, int LS, int NL, int T, bool INV, int MAXR = 8, bool OPQ = false, bool IP = true, class C>
__device__ __forceinline__ void fft_lines_from_pos(C *buf, const C *__restrict__ tw, int twmul)
{
    if constexpr (lines_in_place<C, IP>())
        fft_lines_ip<L, LS, NL, T, INV, true, MAXR, OPQ>(buf, tw, twmul);
    else
        fft_lines<L, LS, NL, T, INV, MAXR, OPQ>(buf, tw, twmul);
}

template <int L, int LS, int NL, int T, bool INV, int S, int MAXR, bool OPQ, class C>
__device__ __forceinline__ void stages_from(C *buf, const C *__restrict__ tw, int twmul)
{
    if constexpr (S < num_stages(L, MAXR)) {
        constexpr int R = stage_radix(L, S, MAXR);
        if constexpr (mfma_kind<C>::value != 0 && R <= 16)
            stage_mfma<L, LS, NL, T, R, stage_span(L, S, MAXR), INV, OPQ, C>(buf, tw, twmul);
        else
            stage<L, LS, NL, T, R, stage_span(L, S, MAXR), INV, OPQ, C>(buf, tw, twmul);
        stages_from<L, LS, NL, T, INV, S + 1, MAXR, OPQ, C>(buf, tw, twmul);
    }
}

// Transform NL lines of length L in place (LDS).  With cm3 / cmh lines every stage of
// radix <= 16 runs on the matrix cores (stage_mfma) and the plan is the radix-16 one.  Element i of line l lives at
// buf[l * LS + pidx(i)]; LS >= padded_len(L).  Callers barrier before (data
// written) -- the first stage reads after the caller's __syncthreads -- and the
// engine ends with a barrier, so results are visible to every thread on return.
// tw may point to global memory or to an LDS table.  MAXR selects the radix plan (8 or 16,
// see pick_radix); OPQ keeps each stage's index arithmetic inside the caller's loops.
template <int L, int LS, int NL, int T, bool INV, int MAXR, bool OPQ, class C>
__device__ __forceinline__ void fft_lines(C *buf, const C *__restrict__ tw, int twmul)
{
    static_assert(smooth7(L), "line length must factor into 2, 3, 5, 7");
    static_assert(LS >= padded_len(L), "line stride must hold the padded line");
    static_assert(MAXR == 8 || MAXR == 16, "radix plans: 8 or 16");
    // matrix-core lines take the radix-16 plan: fewer, larger products
    constexpr int MR = mfma_kind<C>::value != 0 ? 16 : MAXR;
    stages_from<L, LS, NL, T, INV, 0, MR, OPQ, C>(buf, tw, twmul);
}

}  // namespace dnsfft
