/*
 * dns_oracle.c -- FP64 CPU oracle for the 2D pseudo-spectral vorticity step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * What each part restates (reference file:line) and what pins it:
 *   or_fft / Stockham mixed radix   cuFFT's unnormalized C2C (closed library, not in the
 *                                    mount; conventions from vfft_cuda.cu:14-42 and
 *                                    dns_cuda.cu:304,319-371).  Pinned: brute-force DFT,
 *                                    numpy pocketfft golden vectors, Parseval, shift theorem.
 *   or_c2r_2d / or_r2c_2d           2D C2R/R2C plans dns_cuda.cu:319-371.  Pinned: numpy
 *                                    irfft2/rfft2 golden vectors, round trip = M^2 * identity.
 *   or_init_pao                      pao_km3_cuda.cu:38-498.  Pinned: the printed PAO block,
 *                                    PAO seed, [NEXTDT INIT] and it=1 lines of
 *                                    sim_output_{512,4096,16384}_*_gpu.txt.
 *   or_calcom                        calcom_cuda.cu:84-174.  Pinned: om -> (u,w) -> om
 *                                    round trip, closed form on single modes.
 *   or_step2a                        step2a_cuda.cu:59-261.
 *   or_step2b                        step2b_cuda.cu:137-276.
 *   or_step3                         step3_cuda.cu:35-173.  Pinned: single-shell decay
 *                                    factor (1-a)/(1+a), hand-computed two-mode FN,
 *                                    energy/enstrophy conservation of FN.
 *   or_next_dt / or_cflm             dns_cuda.cu:485-580.  Pinned: brute max, KATs.
 *   or_run                           mem_cuda.cu:1760-1789, 1848-1921 (loop order, t update,
 *                                    diagnostic cadence).  Pinned: KAT it=1 line.
 *   or_step_adv                      advective form of the same FN (no reference code;
 *                                    the equivalence is derived).  Pinned: agreement with
 *                                    the uiuj form to roundoff, conservation laws.
 *   or_init_config                   k^4 exp(-2 k^2/K0^2) spectrum (no reference code).
 *                                    Pinned: Hermitian kx=0 column, unit energy, shell
 *                                    spectrum shape, determinism.
 *   or_display_sigma                 mem_cuda.cu:651-870.  Pinned: brute recomputation
 *                                    on small fields, KAT SIG at it=1.
 *   or_pal_over_ens_kmax2            mem_cuda.cu:872-935.  Pinned: single-mode fields give
 *                                    K^2 / kmax2 exactly; two modes on the weight-1 and
 *                                    weight-2 columns give the hand-computed weighted mean;
 *                                    the (0,0) mode leaves the ratio unchanged.
 */
#include "dns_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

static inline cplx cmk(double r, double i) { cplx c = {r, i}; return c; }
static inline cplx cadd(cplx a, cplx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline cplx csub(cplx a, cplx b) { return cmk(a.re - b.re, a.im - b.im); }
static inline cplx cmul(cplx a, cplx b)
{
    return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cplx cscale(cplx a, double s) { return cmk(a.re * s, a.im * s); }
static inline cplx cconj(cplx a) { return cmk(a.re, -a.im); }
/* multiply by sign*i */
static inline cplx cmuli(cplx a, int sign)
{
    return sign > 0 ? cmk(-a.im, a.re) : cmk(a.im, -a.re);
}

/* ===================================================================== */
/* FFT: Stockham autosort, mixed radix 2/3/4/5/7/8 plus generic odd.      */
/* ===================================================================== */

#define OR_MAXST 64
#define OR_MAXR 64

typedef struct {
    int n, nst;
    int rad[OR_MAXST];
    double *twc, *tws;  /* cos/sin(2 pi t / n), t in [0, n) */
} fplan;

static void twiddle_table(int n, double *c, double *s)
{
    for (int t = 0; t < n; ++t) {
        /* Reduce to the first octant so that long double evaluation rounds
         * to the correctly rounded double in practice. */
        long double x = (long double)t / (long double)n;  /* turns in [0,1) */
        long double a = 2.0L * 3.141592653589793238462643383279502884L * x;
        c[t] = (double)cosl(a);
        s[t] = (double)sinl(a);
    }
}

static int fplan_init(fplan *p, int n)
{
    memset(p, 0, sizeof(*p));
    if (n < 1) return -1;
    p->n = n;
    int m = n;
    static const int pref[] = {8, 4, 2, 3, 5, 7};
    while (m > 1) {
        int found = 0;
        for (int i = 0; i < 6; ++i) {
            int r = pref[i];
            /* use 8/4 only while the remainder stays divisible */
            if (m % r == 0) {
                p->rad[p->nst++] = r;
                m /= r;
                found = 1;
                break;
            }
        }
        if (!found) {
            int r = 11;
            while (m % r) r += 2;
            if (r > OR_MAXR) return -2;
            p->rad[p->nst++] = r;
            m /= r;
        }
        if (p->nst >= OR_MAXST) return -3;
    }
    p->twc = (double *)malloc(sizeof(double) * (size_t)n);
    p->tws = (double *)malloc(sizeof(double) * (size_t)n);
    twiddle_table(n, p->twc, p->tws);
    return 0;
}

static void fplan_free(fplan *p)
{
    free(p->twc);
    free(p->tws);
    p->twc = p->tws = NULL;
}

/* constants of the small odd DFTs: cos/sin(2 pi q / R) */
static double g_rc[OR_MAXR + 1][OR_MAXR];
static double g_rs[OR_MAXR + 1][OR_MAXR];
static int g_rinit[OR_MAXR + 1];

static void radix_consts(int R)
{
    if (g_rinit[R]) return;
    for (int q = 0; q < R; ++q) {
        long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)q / (long double)R;
        g_rc[R][q] = (double)cosl(a);
        g_rs[R][q] = (double)sinl(a);
    }
    g_rinit[R] = 1;
}

static void dft_small(cplx *v, int R, int sign)
{
    if (R == 1) return;
    if (R == 2) {
        cplx a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
        return;
    }
    if (R == 4) {
        cplx a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
        cplx s02 = cadd(a0, a2), d02 = csub(a0, a2);
        cplx s13 = cadd(a1, a3), d13 = cmuli(csub(a1, a3), sign);
        v[0] = cadd(s02, s13);
        v[1] = cadd(d02, d13);
        v[2] = csub(s02, s13);
        v[3] = csub(d02, d13);
        return;
    }
    if (R == 8) {
        cplx e[4] = {v[0], v[2], v[4], v[6]};
        cplx o[4] = {v[1], v[3], v[5], v[7]};
        dft_small(e, 4, sign);
        dft_small(o, 4, sign);
        const double h = 0.70710678118654752440084436210484903928;
        /* w8^q = e^{sign 2 pi i q / 8} */
        cplx w1 = cmk(h, sign * h), w3 = cmk(-h, sign * h);
        cplx t0 = o[0];
        cplx t1 = cmul(o[1], w1);
        cplx t2 = cmuli(o[2], sign);
        cplx t3 = cmul(o[3], w3);
        v[0] = cadd(e[0], t0); v[4] = csub(e[0], t0);
        v[1] = cadd(e[1], t1); v[5] = csub(e[1], t1);
        v[2] = cadd(e[2], t2); v[6] = csub(e[2], t2);
        v[3] = cadd(e[3], t3); v[7] = csub(e[3], t3);
        return;
    }
    /* generic odd R: pair terms r and R-r */
    radix_consts(R);
    cplx out[OR_MAXR];
    const int hR = (R - 1) / 2;
    cplx sum = v[0];
    cplx sp[OR_MAXR], dm[OR_MAXR];
    for (int r = 1; r <= hR; ++r) {
        sp[r] = cadd(v[r], v[R - r]);
        dm[r] = csub(v[r], v[R - r]);
        sum = cadd(sum, sp[r]);
    }
    out[0] = sum;
    for (int q = 1; q <= hR; ++q) {
        cplx t = v[0];
        cplx u = cmk(0.0, 0.0);
        for (int r = 1; r <= hR; ++r) {
            int idx = (r * q) % R;
            t = cadd(t, cscale(sp[r], g_rc[R][idx]));
            u = cadd(u, cscale(dm[r], g_rs[R][idx]));
        }
        /* sum_r dm * i*sign*sin */
        cplx iu = cmuli(u, sign);
        out[q] = cadd(t, iu);
        out[R - q] = csub(t, iu);
    }
    for (int q = 0; q < R; ++q) v[q] = out[q];
}

/* x: n complex (in/out); work: n complex scratch */
static void fft_exec(const fplan *p, int sign, cplx *x, cplx *work)
{
    const int n = p->n;
    cplx *src = x, *dst = work;
    int Ns = 1;
    for (int s = 0; s < p->nst; ++s) {
        const int R = p->rad[s];
        const int nR = n / R;
        const int tstride = n / (Ns * R);
        for (int j = 0; j < nR; ++j) {
            const int k = j % Ns;
            cplx v[OR_MAXR];
            for (int r = 0; r < R; ++r) {
                cplx a = src[j + r * nR];
                if (r && k) {
                    int t = k * r * tstride;
                    cplx w = cmk(p->twc[t], sign * p->tws[t]);
                    a = cmul(a, w);
                }
                v[r] = a;
            }
            dft_small(v, R, sign);
            const int idxD = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; ++r) dst[idxD + r * Ns] = v[r];
        }
        cplx *tmp = src; src = dst; dst = tmp;
        Ns *= R;
    }
    if (src != x) memcpy(x, src, sizeof(cplx) * (size_t)n);
}

int or_fft(int n, int sign, const double *in, double *out)
{
    fplan p;
    if (fplan_init(&p, n)) return -1;
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    memcpy(x, in, sizeof(cplx) * (size_t)n);
    fft_exec(&p, sign < 0 ? -1 : 1, x, w);
    memcpy(out, x, sizeof(cplx) * (size_t)n);
    free(x);
    free(w);
    fplan_free(&p);
    return 0;
}

int or_dft_brute(int n, int sign, const double *in, double *out)
{
    const long double tp = 2.0L * 3.141592653589793238462643383279502884L;
    for (int q = 0; q < n; ++q) {
        long double sr = 0.0L, si = 0.0L;
        for (int r = 0; r < n; ++r) {
            long long t = ((long long)r * (long long)q) % n;
            long double a = tp * (long double)t / (long double)n;
            long double c = cosl(a), s = (sign < 0 ? -1.0L : 1.0L) * sinl(a);
            long double xr = in[2 * r], xi = in[2 * r + 1];
            sr += xr * c - xi * s;
            si += xr * s + xi * c;
        }
        out[2 * q] = (double)sr;
        out[2 * q + 1] = (double)si;
    }
    return 0;
}

/* ===================================================================== */
/* 2D transforms on the M x M grid                                        */
/* ===================================================================== */

#define COLB 8

/* In-place column C2C over z for kx in [0, nk) of a [M][nk] array. */
static void columns_c2c(const fplan *p, int M, int nk, cplx *a, int sign)
{
    const int nblk = (nk + COLB - 1) / COLB;
#pragma omp parallel
    {
        cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)M * COLB);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)M);
#pragma omp for schedule(static)
        for (int b = 0; b < nblk; ++b) {
            const int k0 = b * COLB;
            const int kb = (nk - k0) < COLB ? (nk - k0) : COLB;
            for (int z = 0; z < M; ++z)
                for (int c = 0; c < kb; ++c)
                    buf[(size_t)c * M + z] = a[(size_t)z * nk + k0 + c];
            for (int c = 0; c < kb; ++c) fft_exec(p, sign, buf + (size_t)c * M, work);
            for (int z = 0; z < M; ++z)
                for (int c = 0; c < kb; ++c)
                    a[(size_t)z * nk + k0 + c] = buf[(size_t)c * M + z];
        }
        free(buf);
        free(work);
    }
}

static fplan *plan_for(int M);

/* C2R of a [M][nk] half-plane whose content is destroyed (it is the column pass's buffer) */
static void c2r_2d_inplace(int M, cplx *tmp, double *ur)
{
    const int nk = M / 2 + 1;
    const fplan *p = plan_for(M);
    columns_c2c(p, M, nk, tmp, +1);
    const int npairs = (M + 1) / 2;
#pragma omp parallel
    {
        cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)M);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)M);
#pragma omp for schedule(static)
        for (int pr = 0; pr < npairs; ++pr) {
            const int ra = 2 * pr, rb = 2 * pr + 1;
            const cplx *A = tmp + (size_t)ra * nk;
            const cplx *B = (rb < M) ? tmp + (size_t)rb * nk : NULL;
            for (int k = 0; k < M; ++k) {
                cplx a, b;
                if (k == 0 || 2 * k == M) {
                    a = cmk(A[k].re, 0.0);
                    b = B ? cmk(B[k].re, 0.0) : cmk(0.0, 0.0);
                } else if (2 * k < M) {
                    a = A[k];
                    b = B ? B[k] : cmk(0.0, 0.0);
                } else {
                    a = cconj(A[M - k]);
                    b = B ? cconj(B[M - k]) : cmk(0.0, 0.0);
                }
                /* z = a + i b */
                z[k] = cmk(a.re - b.im, a.im + b.re);
            }
            fft_exec(p, +1, z, work);
            double *xa = ur + (size_t)ra * M;
            double *xb = (rb < M) ? ur + (size_t)rb * M : NULL;
            for (int n = 0; n < M; ++n) {
                xa[n] = z[n].re;
                if (xb) xb[n] = z[n].im;
            }
        }
        free(z);
        free(work);
    }
}

void or_c2r_2d(int M, const double *uc_in, double *ur)
{
    const int nk = M / 2 + 1;
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)M * nk);
    memcpy(tmp, uc_in, sizeof(cplx) * (size_t)M * nk);
    c2r_2d_inplace(M, tmp, ur);
    free(tmp);
}

void or_r2c_2d(int M, const double *ur, double *uc_out)
{
    const int nk = M / 2 + 1;
    const fplan *p = plan_for(M);
    cplx *uc = (cplx *)uc_out;
    const int npairs = (M + 1) / 2;
#pragma omp parallel
    {
        cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)M);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)M);
#pragma omp for schedule(static)
        for (int pr = 0; pr < npairs; ++pr) {
            const int ra = 2 * pr, rb = 2 * pr + 1;
            const double *xa = ur + (size_t)ra * M;
            const double *xb = (rb < M) ? ur + (size_t)rb * M : NULL;
            for (int n = 0; n < M; ++n) z[n] = cmk(xa[n], xb ? xb[n] : 0.0);
            fft_exec(p, -1, z, work);
            cplx *A = uc + (size_t)ra * nk;
            cplx *B = (rb < M) ? uc + (size_t)rb * nk : NULL;
            for (int k = 0; k < nk; ++k) {
                cplx zk = z[k];
                cplx zm = cconj(z[(M - k) % M]);
                A[k] = cscale(cadd(zk, zm), 0.5);
                if (B) {
                    cplx d = csub(zk, zm); /* = 2 i B */
                    B[k] = cmk(0.5 * d.im, -0.5 * d.re);
                }
            }
        }
        free(z);
        free(work);
    }
    columns_c2c(p, M, nk, uc, -1);
}

/* plan cache: a handful of sizes per process */
#define OR_NPLANS 16
static fplan g_plans[OR_NPLANS];
static int g_nplans = 0;

static fplan *plan_for(int M)
{
    fplan *found = NULL;
#pragma omp critical(or_plan_cache)
    {
        for (int i = 0; i < g_nplans; ++i)
            if (g_plans[i].n == M) found = &g_plans[i];
        if (!found && g_nplans < OR_NPLANS) {
            fplan_init(&g_plans[g_nplans], M);
            for (int s = 0; s < g_plans[g_nplans].nst; ++s) radix_consts(g_plans[g_nplans].rad[s]);
            found = &g_plans[g_nplans++];
        }
    }
    return found;
}

/* ===================================================================== */
/* Solver state                                                           */
/* ===================================================================== */

struct or_state {
    int N, M, H, NK;
    int it, ifn;
    double Re, K0, visc, t, dt, cn, cnm1, cflnum, cflm;
    cplx *om2, *fnm1;  /* [N][H] */
    cplx *uc;          /* [3][M][NK] */
    double *ur;        /* [3][M][M] */
};

#define UCI(S, kx, z, c) ((size_t)(kx) + (size_t)(S)->NK * ((size_t)(z) + (size_t)(S)->M * (size_t)(c)))
#define URI(S, x, z, c) ((size_t)(x) + (size_t)(S)->M * ((size_t)(z) + (size_t)(S)->M * (size_t)(c)))

or_state *or_create(int N, double Re, double K0, double cflnum)
{
    if (N < 4 || (N % 4) != 0) return NULL;
    or_state *S = (or_state *)calloc(1, sizeof(or_state));
    S->N = N;
    S->M = 3 * N / 2;
    S->H = N / 2;
    S->NK = 3 * N / 4 + 1;
    S->Re = Re;
    S->K0 = K0;
    S->visc = 0.0;
    S->t = 0.0;
    S->dt = 0.0;
    S->cn = 1.0;
    S->cnm1 = 0.0;
    S->cflnum = cflnum;
    S->it = 0;
    S->ifn = 1;
    S->om2 = (cplx *)calloc((size_t)N * S->H, sizeof(cplx));
    S->fnm1 = (cplx *)calloc((size_t)N * S->H, sizeof(cplx));
    S->uc = (cplx *)calloc((size_t)3 * S->M * S->NK, sizeof(cplx));
    S->ur = (double *)calloc((size_t)3 * S->M * S->M, sizeof(double));
    if (!S->om2 || !S->fnm1 || !S->uc || !S->ur) {
        or_destroy(S);
        return NULL;
    }
    plan_for(S->M);
    return S;
}

void or_destroy(or_state *S)
{
    if (!S) return;
    free(S->om2);
    free(S->fnm1);
    free(S->uc);
    free(S->ur);
    free(S);
}

void or_get_scalars(const or_state *S, or_scalars *o)
{
    o->N = S->N; o->M = S->M; o->H = S->H; o->NK = S->NK;
    o->it = S->it; o->ifn = S->ifn;
    o->Re = S->Re; o->K0 = S->K0; o->visc = S->visc; o->t = S->t; o->dt = S->dt;
    o->cn = S->cn; o->cnm1 = S->cnm1; o->cflnum = S->cflnum; o->cflm = S->cflm;
}

void or_set_scalars(or_state *S, const or_scalars *o)
{
    S->it = o->it; S->ifn = o->ifn;
    S->Re = o->Re; S->K0 = o->K0; S->visc = o->visc; S->t = o->t; S->dt = o->dt;
    S->cn = o->cn; S->cnm1 = o->cnm1; S->cflnum = o->cflnum; S->cflm = o->cflm;
}

void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* wavenumbers: alfa = ix, gamma = iz (iz < N/2) else iz - N (pao_km3_cuda.cu:75-92) */
static inline double gam_of(int iz, int N) { return (iz < N / 2) ? (double)iz : (double)(iz - N); }

/* ===================================================================== */
/* PAO-k^-3 initial condition (pao_km3_cuda.cu:38-498), FP32 where the    */
/* reference computes in float.                                           */
/* ===================================================================== */

static inline float pao_frand(int *seed)
{
    const int IMM = 420029, IT = 2017, ID = 5011;
    *seed = (*seed * IMM + IT) % ID;
    return (float)(*seed) / (float)ID;
}

static inline double pao_hash01(int x, int z, int seed, int salt)
{
    const double a = ((double)x + 1.0) * 12.9898 + ((double)z + 1.0) * 78.233 +
                     ((double)seed + 1.0) * 0.010001 + ((double)salt + 1.0) * 37.719;
    const double s = sin(a) * 43758.5453123;
    return s - floor(s);
}

typedef struct { float re, im; } fcplx;

static int logf_append(char *log, int cap, int *pos, const char *fmt, double v)
{
    if (!log || *pos >= cap) return 0;
    int w = snprintf(log + *pos, (size_t)(cap - *pos), fmt, v);
    if (w > 0) *pos += w;
    return w;
}

/* Generates the compact PAO-k^-3 velocity spectrum UR[x][z][c] (x < N/2, z < N,
 * c = u1, u3) as float pairs, prints the diagnostics block into log and returns the
 * final LCG seed.  *out_ur receives the array (caller frees) unless out_ur is NULL. */
static int pao_generate(int N, double Re, double K0d, float **out_ur, char *log, int logcap, double *visc_out)
{
    const int NE = N, ND2 = N / 2, NED2 = NE / 2;
    const float PI = 3.14159265358979f;
    const float DXZ = 2.0f * PI / (float)N;
    const float K0 = (float)K0d;
    const float NORM = PI * K0 * K0;
    const float AMP2_FLOOR = 1.0e-40f;

    float *alfa = (float *)malloc(sizeof(float) * ND2);
    float *gamma = (float *)malloc(sizeof(float) * NE);
    const float E1 = 1.0f, E3 = 1.0f / E1;
    const float DALFA = 1.0f / E1, DGAMMA = 1.0f / E3;
    for (int x = 0; x < NED2; ++x) alfa[x] = (float)x * DALFA;
    gamma[0] = 0.0f;
    for (int z = 1; z <= NED2; ++z) {
        gamma[z] = (float)z * DGAMMA;
        gamma[NE - z] = -gamma[z];
    }

    fcplx *UR = (fcplx *)calloc((size_t)ND2 * NE * 2, sizeof(fcplx));
#define URAT(x, z, c) UR[(size_t)(c) + 2 * ((size_t)(x) + (size_t)ND2 * (size_t)(z))]

    const int seed_init = 1;
    int seed = seed_init;
    float RANVEC[97];
    for (int i = 0; i < 97; ++i) pao_frand(&seed);
    for (int i = 0; i < 97; ++i) RANVEC[i] = pao_frand(&seed);

    for (int z = 0; z < NE; ++z) {
        for (int x = 0; x < NED2; ++x) {
            float r = pao_frand(&seed);
            int idx = (int)(r * 97.0f);
            if (idx < 0) idx = 0;
            if (idx > 96) idx = 96;
            float rv = RANVEC[idx];
            RANVEC[idx] = r;
            float th = 2.0f * PI * rv;
            float ac = cosf(th), as = sinf(th);
            float ax = alfa[x], gz = gamma[z];
            float K = sqrtf(ax * ax + gz * gz);
            if (ax == 0.0f) {
                URAT(x, z, 1).re = 0.0f; URAT(x, z, 1).im = 0.0f;
                float q = K / K0;
                float ABSU2 = expf(-q * q) / NORM;
                if (ABSU2 < AMP2_FLOOR) ABSU2 = AMP2_FLOOR;
                float amp = sqrtf(ABSU2);
                URAT(x, z, 0).re = amp * ac;
                URAT(x, z, 0).im = amp * as;
            } else {
                float denom = 1.0f + (gz * gz) / (ax * ax);
                float q = K / K0;
                float ABSW2 = expf(-q * q) / (denom * NORM);
                if (ABSW2 < AMP2_FLOOR) ABSW2 = AMP2_FLOOR;
                float ampw = sqrtf(ABSW2);
                fcplx w = {ampw * ac, ampw * as};
                float s = -(gz / ax);
                URAT(x, z, 1) = w;
                URAT(x, z, 0).re = s * w.re;
                URAT(x, z, 0).im = s * w.im;
            }
        }
    }
    URAT(0, 0, 0).re = URAT(0, 0, 0).im = 0.0f;
    URAT(0, 0, 1).re = URAT(0, 0, 1).im = 0.0f;
    for (int z = 1; z < NED2; ++z) {
        for (int c = 0; c < 2; ++c) {
            URAT(0, NE - z, c).re = URAT(0, z, c).re;
            URAT(0, NE - z, c).im = -URAT(0, z, c).im;
        }
    }
    for (int x = 0; x < ND2; ++x)
        for (int c = 0; c < 2; ++c) URAT(x, NED2, c).re = URAT(x, NED2, c).im = 0.0f;

    /* k^-3 reshaping (pao_km3_cuda.cu:202-340) */
    {
        const double k_nyquist = (double)ND2;
        const int kmax_shell = ND2;
        double *shell_energy = (double *)calloc((size_t)kmax_shell + 1, sizeof(double));
        double *shell_weight = (double *)calloc((size_t)kmax_shell + 1, sizeof(double));
        double *target_shell = (double *)calloc((size_t)kmax_shell + 1, sizeof(double));
        for (int x = 0; x < ND2; ++x) {
            const double m = (x == 0) ? 1.0 : 2.0;
            const double ax2 = (double)alfa[x] * (double)alfa[x];
            for (int z = 0; z < NE; ++z) {
                const double gz = (double)gamma[z];
                const double gz2 = gz * gz;
                const double k = sqrt(ax2 + gz2);
                const int shell = (int)(k + 0.5);
                const fcplx U1 = URAT(x, z, 0), U3 = URAT(x, z, 1);
                const double u1u1 = (double)U1.re * (double)U1.re + (double)U1.im * (double)U1.im;
                const double u3u3 = (double)U3.re * (double)U3.re + (double)U3.im * (double)U3.im;
                if (k > 0.0 && k <= k_nyquist && shell <= kmax_shell) {
                    int z_hash = z;
                    if (x == 0 && z > NED2) z_hash = NE - z;
                    const double mode_jitter = 0.75 + 0.50 * pao_hash01(x, z_hash, seed_init, 17);
                    shell_energy[shell] += m * (u1u1 + u3u3);
                    shell_weight[shell] += m * mode_jitter;
                }
            }
        }
        const double k0 = (1.0 > (double)K0) ? 1.0 : (double)K0;
        for (int shell = 1; shell <= kmax_shell; ++shell) {
            const double k = (double)shell;
            double shape;
            if (k < k0) {
                const double low = k / k0;
                shape = (low * low * low * low) / (k0 * k0 * k0);
            } else {
                shape = 1.0 / (k * k * k);
            }
            target_shell[shell] = shape;
        }
        double original_energy = 0.0, target_energy = 0.0;
        for (int shell = 1; shell <= kmax_shell; ++shell) {
            original_energy += shell_energy[shell];
            target_energy += target_shell[shell];
        }
        if (original_energy > 0.0 && target_energy > 0.0) {
            const double norm = original_energy / target_energy;
            for (int x = 0; x < ND2; ++x) {
                const double ax = (double)alfa[x];
                const double ax2 = ax * ax;
                for (int z = 0; z < NE; ++z) {
                    const double gz = (double)gamma[z];
                    const double gz2 = gz * gz;
                    const double k = sqrt(ax2 + gz2);
                    const int shell = (int)(k + 0.5);
                    if (k > 0.0 && k <= k_nyquist && shell <= kmax_shell && shell_weight[shell] > 0.0) {
                        int z_hash = z;
                        double phase_sign = 1.0;
                        if (x == 0 && z > NED2) {
                            z_hash = NE - z;
                            phase_sign = -1.0;
                        }
                        const double mode_jitter = 0.75 + 0.50 * pao_hash01(x, z_hash, seed_init, 17);
                        const double mode_energy =
                            (norm * target_shell[shell]) * mode_jitter / shell_weight[shell];
                        const double th =
                            phase_sign * 2.0 * 3.14159265358979323846 * pao_hash01(x, z_hash, seed_init, 53);
                        const float pr = (float)cos(th), pi_ = (float)sin(th);
                        if (ax == 0.0) {
                            const float a = (float)sqrt(mode_energy);
                            URAT(x, z, 0).re = a * pr;
                            URAT(x, z, 0).im = a * pi_;
                            URAT(x, z, 1).re = URAT(x, z, 1).im = 0.0f;
                        } else {
                            const double denom = 1.0 + gz2 / ax2;
                            const float a = (float)sqrt(mode_energy / denom);
                            const fcplx w = {a * pr, a * pi_};
                            const float s = (float)(-gz / ax);
                            URAT(x, z, 0).re = s * w.re;
                            URAT(x, z, 0).im = s * w.im;
                            URAT(x, z, 1) = w;
                        }
                    } else {
                        URAT(x, z, 0).re = URAT(x, z, 0).im = 0.0f;
                        URAT(x, z, 1).re = URAT(x, z, 1).im = 0.0f;
                    }
                }
            }
        }
        free(shell_energy);
        free(shell_weight);
        free(target_shell);
    }

    /* averages (pao_km3_cuda.cu:345-379); std::norm(complex<float>) is a float sum */
    double A1 = 0, A2 = 0, A3 = 0, A4 = 0, A5 = 0, A6 = 0, A7 = 0, E110 = 0;
    for (int x = 0; x < ND2; ++x) {
        const int x1 = (x == 0);
        for (int z = 0; z < N; ++z) {
            const fcplx U1 = URAT(x, z, 0), U3 = URAT(x, z, 1);
            const float n1a = U1.re * U1.re, n1b = U1.im * U1.im;
            const float n3a = U3.re * U3.re, n3b = U3.im * U3.im;
            const float n1 = n1a + n1b, n3 = n3a + n3b;
            const double u1u1 = (double)n1, u3u3 = (double)n3;
            const double ax2 = (double)alfa[x] * (double)alfa[x];
            const double gz2 = (double)gamma[z] * (double)gamma[z];
            const double K2 = ax2 + gz2;
            const double m = x1 ? 1.0 : 2.0;
            A1 += m * u1u1;
            A2 += m * u3u3;
            A3 += m * u1u1 * ax2;
            A4 += m * u1u1 * gz2;
            A5 += m * u3u3 * ax2;
            A6 += m * u3u3 * gz2;
            A7 += m * (u1u1 + u3u3) * K2 * K2;
            if (x1) E110 += u1u1;
        }
    }
    const double Q2 = A1 + A2;
    const double W2 = A3 + A4 + A5 + A6;
    const double visc = 1.0 / Re;
    *visc_out = visc;
    const double EP = visc * W2;
    const double De = 2.0 * visc * visc * A7;
    const double KOL = pow(visc * visc * visc / EP, 0.25);
    double NLAM = 0.0;
    if (E110 != 0.0) NLAM = 2.0 * A1 / E110;
    const double a11 = 2.0 * A1 / Q2 - 1.0;
    const double e11 = 2.0 * (A3 + A4) / W2 - 1.0;
    const double tscale = 0.5 * Q2 / EP;
    const double dxKol = (double)DXZ / KOL;
    const double Lux = 2.0 * (double)PI / sqrt(2.0 * A1 / A3);
    const double Luz = 2.0 * (double)PI / sqrt(2.0 * A1 / A4);
    const double Lwx = 2.0 * (double)PI / sqrt(2.0 * A2 / A5);
    const double Lwz = 2.0 * (double)PI / sqrt(2.0 * A2 / A6);
    const double Ceps2 = 0.5 * Q2 * De / (EP * EP);

    int pos = 0;
    if (log && logcap > 0) log[0] = '\0';
    logf_append(log, logcap, &pos, " Reynolds n. =%12.1f\n", Re);
    logf_append(log, logcap, &pos, " Energy      =%12.7f\n", Q2);
    logf_append(log, logcap, &pos, " WiWi        =%12.5f\n", W2);
    logf_append(log, logcap, &pos, " Epsilon     =%12.7f\n", EP);
    logf_append(log, logcap, &pos, " a11         =%12.7f\n", a11);
    logf_append(log, logcap, &pos, " e11         =%12.7f\n", e11);
    logf_append(log, logcap, &pos, " Time scale  =%12.4f\n", tscale);
    logf_append(log, logcap, &pos, " Kolmogorov  =%12.7f\n", KOL);
    logf_append(log, logcap, &pos, " Viscosity   =%12.7f\n", visc);
    logf_append(log, logcap, &pos, " dx/Kol.     =%12.7f\n", dxKol);
    logf_append(log, logcap, &pos, " 2Pi/Nlamda  =%12.7f\n", NLAM);
    logf_append(log, logcap, &pos, " 2Pi/Lux     =%12.7f\n", Lux);
    logf_append(log, logcap, &pos, " 2Pi/Luz     =%12.7f\n", Luz);
    logf_append(log, logcap, &pos, " 2Pi/Lwx     =%12.7f\n", Lwx);
    logf_append(log, logcap, &pos, " 2Pi/Lwz     =%12.7f\n", Lwz);
    logf_append(log, logcap, &pos, " Deps.       =%12.7f\n", De);
    logf_append(log, logcap, &pos, " Ceps2       =%12.7f\n", Ceps2);
    logf_append(log, logcap, &pos, " E1          =%12.7f\n", (double)E1);
    logf_append(log, logcap, &pos, " E3          =%12.7f\n", (double)E3);
    if (log && pos < logcap) {
        int w = snprintf(log + pos, (size_t)(logcap - pos), " PAO seed    =%12d\n", seed);
        if (w > 0) pos += w;
    }

#undef URAT
    free(alfa);
    free(gamma);
    if (out_ur)
        *out_ur = (float *)UR;
    else
        free(UR);
    return seed;
}

int or_pao_diagnostics(int N, double Re, double K0, char *log, int logcap)
{
    double visc;
    return pao_generate(N, Re, K0, NULL, log, logcap, &visc);
}

int or_init_pao(or_state *S, char *log, int logcap)
{
    float *ur = NULL;
    const int N = S->N, ND2 = N / 2;
    const int seed = pao_generate(N, S->Re, S->K0, &ur, log, logcap, &S->visc);
    /* scatter compact [x][z][c] into UC rows 0..N-1 (k_scatter_pao_uc_to_full) */
    memset(S->uc, 0, sizeof(cplx) * (size_t)3 * S->M * S->NK);
    for (int z = 0; z < N; ++z)
        for (int x = 0; x < ND2; ++x)
            for (int c = 0; c < 2; ++c) {
                const float *v = ur + 2 * ((size_t)c + 2 * ((size_t)x + (size_t)ND2 * (size_t)z));
                S->uc[UCI(S, x, z, c)] = cmk((double)v[0], (double)v[1]);
            }
    free(ur);
    or_calcom(S);
    return seed;
}

/* CALCOM: om2 = i (gamma*u1 - alfa*u2) from UC rows 0..N-1 (calcom_cuda.cu:84-132) */
void or_calcom(or_state *S)
{
    const int N = S->N, H = S->H;
    for (int iz = 0; iz < N; ++iz) {
        const double gz = gam_of(iz, N);
        for (int ix = 0; ix < H; ++ix) {
            const double ax = (double)ix;
            const cplx u1 = S->uc[UCI(S, ix, iz, 0)];
            const cplx u2 = S->uc[UCI(S, ix, iz, 1)];
            const cplx d = cmk(gz * u1.re - ax * u2.re, gz * u1.im - ax * u2.im);
            S->om2[(size_t)iz * H + ix] = cmk(-d.im, d.re);
        }
    }
}

/* velocity spectrum from omega (step3_cuda.cu:102-138) */
static inline void vel_from_om(cplx om, int ix, int iz, int N, cplx *u, cplx *w)
{
    const double ax = (double)ix;
    const double gz = gam_of(iz, N);
    if (ix >= 1) {
        const double K2 = ax * ax + gz * gz;
        const double invK2 = 1.0 / (K2 + 1.0e-30);
        const cplx v = cscale(om, invK2);
        *u = cmk(gz * v.im, -gz * v.re);
        *w = cmk(-ax * v.im, ax * v.re);
    } else {
        const double invG = (iz >= 1 && fabs(gz) > 0.0) ? (1.0 / gz) : 0.0;
        const cplx v = cscale(om, invG);
        *u = cmk(v.im, -v.re);
        *w = cmk(0.0, 0.0);
    }
}

/* ---- config IC ------------------------------------------------------ */

static inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

double or_ic_uniform(uint64_t seed, int kx, int kz)
{
    const uint64_t key = ((uint64_t)(uint32_t)kx << 32) | (uint64_t)(uint32_t)kz;
    const uint64_t h = splitmix64(seed ^ splitmix64(key));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

static inline double cfg_mode_energy(double K, double K0)
{
    const double q = K / K0;
    const double q2 = q * q;
    return (q2 * q2) * exp(-2.0 * q2) / (2.0 * 3.14159265358979323846 * K);
}

void or_init_config(or_state *S, uint64_t seed)
{
    const int N = S->N, H = S->H;
    const double K0 = S->K0;
    /* normalization: sum over retained modes (iz ascending, per-row partials) */
    double *rows = (double *)calloc((size_t)N, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < N; ++iz) {
        if (iz == N / 2) continue;
        const double gz = gam_of(iz, N);
        double acc = 0.0;
        for (int ix = 0; ix < H; ++ix) {
            if (ix == 0 && iz == 0) continue;
            const double K = sqrt((double)ix * ix + gz * gz);
            const double m = (ix == 0) ? 1.0 : 2.0;
            acc += m * cfg_mode_energy(K, K0);
        }
        rows[iz] = acc;
    }
    double tot = 0.0;
    for (int iz = 0; iz < N; ++iz) tot += rows[iz];
    free(rows);
    const double c = 1.0 / tot;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < N; ++iz) {
        const double gz = gam_of(iz, N);
        for (int ix = 0; ix < H; ++ix) {
            cplx v = cmk(0.0, 0.0);
            if (!(ix == 0 && iz == 0) && iz != N / 2) {
                int kz = (int)gz;
                int conj = 0;
                if (ix == 0 && kz < 0) {
                    kz = -kz;
                    conj = 1;
                }
                const double K = sqrt((double)ix * ix + (double)kz * kz);
                const double amp = K * sqrt(c * cfg_mode_energy(K, K0));
                const double th = 2.0 * 3.14159265358979323846 * or_ic_uniform(seed, ix, kz);
                v = cmk(amp * cos(th), amp * sin(th));
                if (conj) v.im = -v.im;
            }
            S->om2[(size_t)iz * H + ix] = v;
        }
    }
    S->visc = 1.0 / S->Re;
    /* velocity spectrum into UC rows 0..N-1, like the state after CALCOM */
    memset(S->uc, 0, sizeof(cplx) * (size_t)3 * S->M * S->NK);
    for (int iz = 0; iz < N; ++iz)
        for (int ix = 0; ix < H; ++ix) {
            cplx u, w;
            vel_from_om(S->om2[(size_t)iz * H + ix], ix, iz, N, &u, &w);
            S->uc[UCI(S, ix, iz, 0)] = u;
            S->uc[UCI(S, ix, iz, 1)] = w;
        }
}

/* ===================================================================== */
/* Reference steps                                                        */
/* ===================================================================== */

/* STEP2A: zero kx in [N/2, 3N/4] for comps 0,1; move rows [N/2,N) to [N,3N/2);
 * inverse 2D C2R of comps 0,1 (step2a_cuda.cu:59-82, 135-161, 211-261). */
void or_step2a(or_state *S)
{
    const int N = S->N, M = S->M, NK = S->NK;
    for (int c = 0; c < 2; ++c) {
        for (int z = 0; z < M; ++z)
            for (int kx = N / 2; kx <= 3 * N / 4 && kx < NK; ++kx) S->uc[UCI(S, kx, z, c)] = cmk(0, 0);
        for (int tz = 0; tz < N / 2; ++tz) {
            const int zm = tz + N / 2, zt = tz + N;
            if (zt >= M) continue;
            for (int tx = 0; tx < N / 2 && tx < NK; ++tx) {
                S->uc[UCI(S, tx, zt, c)] = S->uc[UCI(S, tx, zm, c)];
                S->uc[UCI(S, tx, zm, c)] = cmk(0, 0);
            }
        }
    }
    for (int c = 0; c < 2; ++c)
        or_c2r_2d(M, (const double *)(S->uc + (size_t)c * M * NK), S->ur + (size_t)c * M * M);
}

/* STEP2B: (u,w) -> (u^2, w^2, uw); forward 2D R2C of the 3 comps; zero row z=N
 * (the kz=-N/2 row) for kx < N/2 (step2b_cuda.cu:137-173, 203-276). */
void or_step2b(or_state *S)
{
    const int N = S->N, M = S->M, NK = S->NK;
    const size_t plane = (size_t)M * M;
    double *u = S->ur, *w = S->ur + plane, *p = S->ur + 2 * plane;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < plane; ++i) {
        const double a = u[i], b = w[i];
        p[i] = a * b;
        u[i] = a * a;
        w[i] = b * b;
    }
    for (int c = 0; c < 3; ++c)
        or_r2c_2d(M, S->ur + (size_t)c * plane, (double *)(S->uc + (size_t)c * M * NK));
    for (int c = 0; c < 3; ++c)
        for (int kx = 0; kx < N / 2; ++kx) S->uc[UCI(S, kx, N, c)] = cmk(0, 0);
}

/* STEP3: FN from the uiuj spectra, CN/AB2 update, velocity from the new omega
 * (step3_cuda.cu:35-173). */
void or_step3(or_state *S)
{
    const int N = S->N, H = S->H, M = S->M;
    const double divxz = 1.0 / ((double)M * (double)M);
    const double VT = 0.5 * S->visc * S->dt;
    const double c2 = 0.5 * S->dt * (2.0 + S->cnm1);
    const double c3 = -0.5 * S->dt * S->cnm1;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < N; ++iz) {
        const double gz = gam_of(iz, N);
        const int z_spec = (iz < N / 2) ? iz : iz + N / 2;
        for (int ix = 0; ix < H; ++ix) {
            const double ax = (double)ix;
            const double A2 = ax * ax, G2 = gz * gz, K2 = A2 + G2;
            const double GA = gz * ax, GmA = G2 - A2;
            const cplx u1 = S->uc[UCI(S, ix, z_spec, 0)];
            const cplx u2 = S->uc[UCI(S, ix, z_spec, 1)];
            const cplx u3 = S->uc[UCI(S, ix, z_spec, 2)];
            const cplx FN = cmk((GA * (u1.re - u2.re) + GmA * u3.re) * divxz,
                                (GA * (u1.im - u2.im) + GmA * u3.im) * divxz);
            const size_t io = (size_t)iz * H + ix;
            const cplx om = S->om2[io], fo = S->fnm1[io];
            const double ARG = VT * K2;
            const double DEN = 1.0 + ARG;
            const double c1 = 1.0 - ARG;
            const cplx num = cmk(c1 * om.re + c2 * FN.re + c3 * fo.re, c1 * om.im + c2 * FN.im + c3 * fo.im);
            const cplx on = cmk(num.re / DEN, num.im / DEN);
            S->om2[io] = on;
            S->fnm1[io] = FN;
            cplx uu, ww;
            vel_from_om(on, ix, iz, N, &uu, &ww);
            S->uc[UCI(S, ix, iz, 0)] = uu;
            S->uc[UCI(S, ix, iz, 1)] = ww;
        }
    }
    S->cnm1 = S->cn;
}

double or_cflm(const or_state *S)
{
    const double PI = 3.14159265358979323846;
    const double DX = 2.0 * PI / (double)S->N, DZ = 2.0 * PI / (double)S->N;
    const size_t plane = (size_t)S->M * S->M;
    const double *u = S->ur, *w = S->ur + plane;
    double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (size_t i = 0; i < plane; ++i) {
        const double a = fabs(u[i]) / DX + fabs(w[i]) / DZ;
        if (a > m) m = a;
    }
    return m;
}

/* NEXTDT (dns_cuda.cu:553-580) */
void or_next_dt(or_state *S)
{
    const double PI = 3.14159265358979323846;
    if (S->it == 0 && S->dt == 0.0) {
        const double CFLM = or_cflm(S);
        S->dt = S->cflnum / CFLM / PI;
        S->cflm = CFLM;
        return;
    }
    if (S->ifn > 0 && (S->it % S->ifn) == 0) {
        const double CFLM = or_cflm(S);
        const double CFL = CFLM * S->dt * PI;
        S->cn = 0.8 + 0.2 * (S->cflnum / CFL);
        S->dt = S->dt * S->cn;
        S->cflm = CFLM;
    }
}

/* ===================================================================== */
/* Advective form                                                         */
/* ===================================================================== */

/* scatter a compact [N][H] spectrum into a zeroed [M][NK] 3/2-grid plane */
static void scatter_compact(const or_state *S, const cplx *src, cplx *dst)
{
    const int N = S->N, H = S->H, NK = S->NK, M = S->M;
    memset(dst, 0, sizeof(cplx) * (size_t)M * NK);
    for (int iz = 0; iz < N; ++iz) {
        const int z = (iz < N / 2) ? iz : iz + N / 2;
        memcpy(dst + (size_t)z * NK, src + (size_t)iz * H, sizeof(cplx) * (size_t)H);
    }
}

void or_velocity_from_om2(or_state *S)
{
    const int N = S->N, H = S->H, M = S->M, NK = S->NK;
    cplx *cu = (cplx *)malloc(sizeof(cplx) * (size_t)N * H);
    cplx *cw = (cplx *)malloc(sizeof(cplx) * (size_t)N * H);
    for (int iz = 0; iz < N; ++iz)
        for (int ix = 0; ix < H; ++ix)
            vel_from_om(S->om2[(size_t)iz * H + ix], ix, iz, N, &cu[(size_t)iz * H + ix], &cw[(size_t)iz * H + ix]);
    cplx *g = (cplx *)malloc(sizeof(cplx) * (size_t)M * NK);
    scatter_compact(S, cu, g);
    c2r_2d_inplace(M, g, S->ur);
    scatter_compact(S, cw, g);
    c2r_2d_inplace(M, g, S->ur + (size_t)M * M);
    free(g);
    free(cu);
    free(cw);
}

/* The advective FN = -FFT(u d_x omega + w d_z omega)/M^2 on the dealiased set.  Memory-lean
 * (it also runs at N = 29400, where one M x M plane is 15.6 GB): each field's spectrum is
 * built straight into the 3/2-grid buffer, which the in-place C2R consumes; the product
 * P = u d_x omega + w d_z omega is formed in two passes over three planes, P = u ox, then
 * P = P + w oz -- with -ffp-contract=off that is the same two roundings of each product and
 * one of the sum as evaluating the expression at once. */
static void field_spec_compact(const or_state *S, int f, cplx *dst)
{
    const int N = S->N, H = S->H;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < N; ++iz) {
        const double gz = gam_of(iz, N);
        for (int ix = 0; ix < H; ++ix) {
            const size_t i = (size_t)iz * H + ix;
            const cplx om = S->om2[i];
            cplx u, w;
            switch (f) {
                case 0: vel_from_om(om, ix, iz, N, &u, &w); dst[i] = u; break;
                case 1: vel_from_om(om, ix, iz, N, &u, &w); dst[i] = w; break;
                case 2: dst[i] = cmk(-(double)ix * om.im, (double)ix * om.re); break; /* i alfa om  */
                default: dst[i] = cmk(-gz * om.im, gz * om.re); break;             /* i gamma om */
            }
        }
    }
}

void or_nonlinear_adv(or_state *S, double *fn_out)
{
    const int N = S->N, H = S->H, M = S->M, NK = S->NK;
    const size_t nc = (size_t)N * H, plane = (size_t)M * M;
    cplx *spec = (cplx *)malloc(sizeof(cplx) * nc);
    double *P = (double *)malloc(sizeof(double) * plane);
    double *A = (double *)malloc(sizeof(double) * plane);
    double *B = (double *)malloc(sizeof(double) * plane);
    cplx *g = (cplx *)malloc(sizeof(cplx) * (size_t)M * NK);
    /* fields 0 (u), 1 (w), 2 (d_x omega), 3 (d_z omega) */
    field_spec_compact(S, 0, spec);
    scatter_compact(S, spec, g);
    c2r_2d_inplace(M, g, A);
    field_spec_compact(S, 2, spec);
    scatter_compact(S, spec, g);
    c2r_2d_inplace(M, g, B);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < plane; ++i) P[i] = A[i] * B[i];
    field_spec_compact(S, 1, spec);
    scatter_compact(S, spec, g);
    c2r_2d_inplace(M, g, A);
    field_spec_compact(S, 3, spec);
    scatter_compact(S, spec, g);
    c2r_2d_inplace(M, g, B);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < plane; ++i) P[i] = P[i] + A[i] * B[i];
    free(A);
    free(B);
    or_r2c_2d(M, P, (double *)g);
    free(P);
    const double divxz = 1.0 / ((double)M * (double)M);
    cplx *fn = (cplx *)fn_out;
    for (int iz = 0; iz < N; ++iz) {
        const int z_spec = (iz < N / 2) ? iz : iz + N / 2;
        for (int ix = 0; ix < H; ++ix) {
            cplx v = cmk(0.0, 0.0);
            if (iz != N / 2) {
                const cplx p = g[(size_t)z_spec * NK + ix];
                v = cmk(-p.re * divxz, -p.im * divxz);
            }
            fn[(size_t)iz * H + ix] = v;
        }
    }
    free(g);
    free(spec);
}

void or_nonlinear_uiuj(or_state *S, double *fn_out)
{
    const int N = S->N, H = S->H, M = S->M, NK = S->NK;
    /* velocity spectrum at compact rows, then the reference 2A -> 2B path */
    cplx *save_uc = (cplx *)malloc(sizeof(cplx) * (size_t)3 * M * NK);
    memcpy(save_uc, S->uc, sizeof(cplx) * (size_t)3 * M * NK);
    memset(S->uc, 0, sizeof(cplx) * (size_t)3 * M * NK);
    for (int iz = 0; iz < N; ++iz)
        for (int ix = 0; ix < H; ++ix)
            vel_from_om(S->om2[(size_t)iz * H + ix], ix, iz, N, &S->uc[UCI(S, ix, iz, 0)], &S->uc[UCI(S, ix, iz, 1)]);
    or_step2a(S);
    or_step2b(S);
    const double divxz = 1.0 / ((double)M * (double)M);
    cplx *fn = (cplx *)fn_out;
    for (int iz = 0; iz < N; ++iz) {
        const double gz = gam_of(iz, N);
        const int z_spec = (iz < N / 2) ? iz : iz + N / 2;
        for (int ix = 0; ix < H; ++ix) {
            const double ax = (double)ix;
            const double GA = gz * ax, GmA = gz * gz - ax * ax;
            const cplx u1 = S->uc[UCI(S, ix, z_spec, 0)];
            const cplx u2 = S->uc[UCI(S, ix, z_spec, 1)];
            const cplx u3 = S->uc[UCI(S, ix, z_spec, 2)];
            fn[(size_t)iz * H + ix] = cmk((GA * (u1.re - u2.re) + GmA * u3.re) * divxz,
                                          (GA * (u1.im - u2.im) + GmA * u3.im) * divxz);
        }
    }
    memcpy(S->uc, save_uc, sizeof(cplx) * (size_t)3 * M * NK);
    free(save_uc);
}

void or_step_adv(or_state *S)
{
    const int N = S->N, H = S->H;
    const size_t nc = (size_t)N * H;
    cplx *FNa = (cplx *)malloc(sizeof(cplx) * nc);
    or_nonlinear_adv(S, (double *)FNa);
    const double VT = 0.5 * S->visc * S->dt;
    const double c2 = 0.5 * S->dt * (2.0 + S->cnm1);
    const double c3 = -0.5 * S->dt * S->cnm1;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < N; ++iz) {
        const double gz = gam_of(iz, N);
        for (int ix = 0; ix < H; ++ix) {
            const double ax = (double)ix;
            const double K2 = ax * ax + gz * gz;
            const size_t io = (size_t)iz * H + ix;
            const cplx FN = FNa[io];
            const cplx om = S->om2[io], fo = S->fnm1[io];
            const double ARG = VT * K2;
            const double DEN = 1.0 + ARG;
            const double c1 = 1.0 - ARG;
            const cplx num = cmk(c1 * om.re + c2 * FN.re + c3 * fo.re, c1 * om.im + c2 * FN.im + c3 * fo.im);
            S->om2[io] = cmk(num.re / DEN, num.im / DEN);
            S->fnm1[io] = FN;
        }
    }
    free(FNa);
    S->cnm1 = S->cn;
}

/* ===================================================================== */
/* Driver loop                                                            */
/* ===================================================================== */

int or_start(or_state *S, int form)
{
    if (form == 0)
        or_step2a(S);
    else
        or_velocity_from_om2(S);
    or_next_dt(S);
    return 0;
}

void or_one_step(or_state *S, int form)
{
    if (form == 0) {
        or_step2b(S);
        or_step3(S);
        or_step2a(S);
    } else {
        or_step_adv(S);
    }
}

static double wall_seconds(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
#endif
}

int or_run(or_state *S, int steps, int update, int form, double *series, int maxrows)
{
    return or_run_timed(S, steps, update, form, series, maxrows, NULL);
}

int or_run_timed(or_state *S, int steps, int update, int form, double *series, int maxrows, double *step_seconds)
{
    int rows = 0;
    if (update <= 0) update = 100;
    const int it0 = S->it;
    for (int it = S->it + 1; it <= steps; ++it) {
        S->it = it;
        const double dt_old = S->dt;
        const double w0 = step_seconds ? wall_seconds() : 0.0;
        or_one_step(S, form);
        if (step_seconds) step_seconds[it - it0 - 1] = wall_seconds() - w0;
        S->t += dt_old;
        if (it == 1 || (it % update) == 0 || it == steps) {
            if (form != 0) or_velocity_from_om2(S);
            or_next_dt(S);
            if (series && rows < maxrows) {
                double E, Z;
                or_energy_enstrophy(S, &E, &Z);
                double *r = series + (size_t)7 * rows;
                r[0] = it; r[1] = S->t; r[2] = S->dt; r[3] = S->cn; r[4] = S->cflm; r[5] = E; r[6] = Z;
                ++rows;
            }
        }
    }
    return rows;
}

/* ===================================================================== */
/* Diagnostics                                                            */
/* ===================================================================== */

void or_energy_enstrophy(const or_state *S, double *E, double *Z)
{
    const int N = S->N, H = S->H;
    double e = 0.0, z = 0.0;
    for (int iz = 0; iz < N; ++iz) {
        const double gz = gam_of(iz, N);
        for (int ix = 0; ix < H; ++ix) {
            const cplx o = S->om2[(size_t)iz * H + ix];
            const double m = (ix == 0) ? 1.0 : 2.0;
            const double a2 = o.re * o.re + o.im * o.im;
            const double K2 = (double)ix * ix + gz * gz;
            z += m * a2;
            if (K2 > 0.0) e += m * a2 / K2;
        }
    }
    *E = e;
    *Z = z;
}

/* Palinstrophy over enstrophy, scaled by the largest K^2 (mem_cuda.cu:872-935, the
 * ADAPT_VISC controller's input): sum w K^2 |om|^2 / ((sum w |om|^2 - |om(0,0)|^2) kmax2)
 * with w = 1 on the ix = 0 and ix = H - 1 columns and 2 elsewhere (the reference's
 * weights, :914-917), kmax2 = max K^2 over the stored modes, 0 when a denominator is not
 * positive (:930-932). */
double or_pal_over_ens_kmax2(const or_state *S)
{
    const int N = S->N, H = S->H;
    double total_full = 0.0, pal = 0.0, kmax2 = 0.0, k00 = 0.0;
    for (int iz = 0; iz < N; ++iz) {
        const double gz = gam_of(iz, N);
        for (int ix = 0; ix < H; ++ix) {
            const cplx o = S->om2[(size_t)iz * H + ix];
            const double w = (ix == 0 || ix == H - 1) ? 1.0 : 2.0;
            const double k2 = (double)ix * ix + gz * gz;
            const double p = o.re * o.re + o.im * o.im;
            total_full += w * p;
            pal += w * k2 * p;
            if (k2 > kmax2) kmax2 = k2;
            if (ix == 0 && iz == 0) k00 = p;
        }
    }
    const double total = total_full - k00;
    if (total <= 0.0 || kmax2 <= 0.0) return 0.0;
    return pal / (total * kmax2);
}

/* OM2PHYS: om2 into its 3/2-grid rows of UC comp 2, inverse to UR comp 2
 * (om2phys_cuda.cu:44-60, 85-109, 153-205). */
void or_om2phys(or_state *S)
{
    const int M = S->M, NK = S->NK;
    cplx *g = S->uc + (size_t)2 * M * NK;
    scatter_compact(S, S->om2, g);
    or_c2r_2d(M, (const double *)g, S->ur + (size_t)2 * M * M);
}

/* display sigma of a physical plane (mem_cuda.cu:651-870), in FP64 */
double or_display_sigma(const or_state *S, int comp)
{
    const size_t plane = (size_t)S->M * S->M;
    const double *f = S->ur + plane * (size_t)comp;
    double mn = f[0], mx = f[0];
    for (size_t i = 0; i < plane; ++i) {
        if (f[i] < mn) mn = f[i];
        if (f[i] > mx) mx = f[i];
    }
    const double rng = mx - mn;
    unsigned long long s = 0, s2 = 0;
    if (fabs(rng) <= 1.0e-12) {
        s = 128ull * plane;
        s2 = 128ull * 128ull * plane;
    } else {
        for (size_t i = 0; i < plane; ++i) {
            double pf = 1.0 + ((f[i] - mn) / rng) * 254.0;
            if (pf < 1.0) pf = 1.0;
            if (pf > 255.0) pf = 255.0;
            const unsigned long long p = (unsigned long long)(int)pf;
            s += p;
            s2 += p * p;
        }
    }
    const double mean = (double)s / (double)plane;
    double var = (double)s2 / (double)plane - mean * mean;
    if (var < 0.0) var = 0.0;
    return sqrt(var);
}

/* ---- raw access ---- */
void or_get_om2(const or_state *S, double *o) { memcpy(o, S->om2, sizeof(cplx) * (size_t)S->N * S->H); }
void or_set_om2(or_state *S, const double *i) { memcpy(S->om2, i, sizeof(cplx) * (size_t)S->N * S->H); }
void or_get_fnm1(const or_state *S, double *o) { memcpy(o, S->fnm1, sizeof(cplx) * (size_t)S->N * S->H); }
void or_set_fnm1(or_state *S, const double *i) { memcpy(S->fnm1, i, sizeof(cplx) * (size_t)S->N * S->H); }
void or_get_ur(const or_state *S, int c, double *o)
{
    memcpy(o, S->ur + (size_t)c * S->M * S->M, sizeof(double) * (size_t)S->M * S->M);
}
void or_get_uc(const or_state *S, int c, double *o)
{
    memcpy(o, S->uc + (size_t)c * S->M * S->NK, sizeof(cplx) * (size_t)S->M * S->NK);
}
void or_set_uc(or_state *S, int c, const double *i)
{
    memcpy(S->uc + (size_t)c * S->M * S->NK, i, sizeof(cplx) * (size_t)S->M * S->NK);
}
