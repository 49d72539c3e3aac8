/*
 * dns_oracle.h -- FP64 CPU restatement of the mem_cuda time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle for the
 * HIP product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  The product
 * (optimizec_amd/, include/dns_hip.h) never links or calls it.
 *
 * Layouts follow the reference (cuda/cuda_dns.h:29-34, 39-103):
 *   om2, fnm1 : [iz][ix]      N rows x H=N/2 complex   (iz = compact kz index)
 *   uc        : [c][z][kx]    3 x M x NK complex       (M = 3N/2, NK = 3N/4+1)
 *   ur        : [c][z][x]     3 x M x M real
 * Complex values are interleaved (re, im) doubles.
 *
 * Parity status per function is recorded in the header of dns_oracle.c.
 */
#ifndef DNS_ORACLE_H
#define DNS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1D transforms -------------------------------------------------- */
/* Unnormalized complex DFT of length n: out[q] = sum_r in[r] e^{sign*2*pi*i*r*q/n}.
 * sign = -1 is the forward transform.  Returns 0 on success. */
int or_fft(int n, int sign, const double *in, double *out);
/* Same transform by the O(n^2) definition in long double (test reference). */
int or_dft_brute(int n, int sign, const double *in, double *out);

/* ---- 2D transforms on the M x M grid (reference conventions) -------- */
/* uc [M][M/2+1] complex -> ur [M][M] real; inverse z (C2C) then inverse x (C2R),
 * unnormalized, kx=0 and kx=M/2 bins taken by their real part. */
void or_c2r_2d(int M, const double *uc, double *ur);
/* ur [M][M] real -> uc [M][M/2+1] complex; forward x (R2C) then forward z (C2C). */
void or_r2c_2d(int M, const double *ur, double *uc);

/* ---- solver state ---------------------------------------------------- */
typedef struct or_state or_state;

typedef struct {
    int    N, M, H, NK;
    int    it, ifn;
    double Re, K0, visc, t, dt, cn, cnm1, cflnum, cflm;
} or_scalars;

or_state *or_create(int N, double Re, double K0, double cflnum);
void      or_destroy(or_state *S);
void      or_get_scalars(const or_state *S, or_scalars *out);
void      or_set_scalars(or_state *S, const or_scalars *in);

/* Initial conditions.  Both leave UC (velocity spectrum, rows 0..N-1) and om2 set,
 * like dnsCudaInit (dns_cuda.cu:591-607) = PAO + CALCOM. */
/* PAO-k^-3 (pao_km3_cuda.cu:59-498).  Writes the printed diagnostics block into
 * log (may be NULL) and returns the final LCG seed. */
int  or_init_pao(or_state *S, char *log, int logcap);
/* The same generator without a solver state (diagnostics block + final seed only). */
int  or_pao_diagnostics(int N, double Re, double K0, char *log, int logcap);
/* Seeded random-phase spectrum E(k) ~ (k/K0)^4 exp(-2 (k/K0)^2), unit energy. */
void or_init_config(or_state *S, uint64_t seed);
/* Counter-based uniform variate in [0,1) keyed by (seed, kx, kz). */
double or_ic_uniform(uint64_t seed, int kx, int kz);

/* Reference steps (uiuj stress form). */
void or_step2a(or_state *S);   /* step2a_cuda.cu:211-261 */
void or_step2b(or_state *S);   /* step2b_cuda.cu:232-276 */
void or_step3(or_state *S);    /* step3_cuda.cu:35-173   */
void or_next_dt(or_state *S);  /* dns_cuda.cu:553-580    */
void or_calcom(or_state *S);   /* calcom_cuda.cu:84-174  */

/* Advective-form step: FN = -FFT(u d_x w + w d_z w)/M^2, same CN/AB2 update. */
void or_step_adv(or_state *S);
/* Physical velocity of the current om2 into ur comps 0,1 (for CFL in adv mode). */
void or_velocity_from_om2(or_state *S);

/* Full driver loop (mem_cuda.cu:1760-1789, 1848-1921):
 *   form 0 = reference uiuj, 1 = advective.
 * Runs steps it+1..steps; on diagnostic steps (it==1, it%update==0, it==steps)
 * calls next_dt and appends (it, t, dt, cn, cflm, E, Z) to series (7 doubles per
 * row) if series != NULL, up to maxrows rows.  Returns number of rows written. */
int  or_start(or_state *S, int form);   /* initial step2a + next_dt (it=0) */
int  or_run(or_state *S, int steps, int update, int form, double *series, int maxrows);
/* or_run, and the wall time of each step's loop body (or_one_step, without the
 * diagnostics) in step_seconds[0 .. steps - it0 - 1] when it is not NULL. */
int  or_run_timed(or_state *S, int steps, int update, int form, double *series, int maxrows,
                  double *step_seconds);
/* One step of the loop body, without diagnostics. */
void or_one_step(or_state *S, int form);

/* Diagnostics. */
void   or_energy_enstrophy(const or_state *S, double *E, double *Z);
double or_cflm(const or_state *S);              /* max |u|/DX + |w|/DZ on ur   */
void   or_om2phys(or_state *S);                 /* om2 -> ur comp 2 (om2phys_cuda.cu:153-205) */
double or_display_sigma(const or_state *S, int comp); /* mem_cuda.cu:820-870 */
/* sum w K^2 |om|^2 / ((sum w |om|^2 - |om(0,0)|^2) max K^2): the ADAPT_VISC input
 * (mem_cuda.cu:872-935) */
double or_pal_over_ens_kmax2(const or_state *S);
/* FN of the current state (advective form) into fn [N][H]. */
void   or_nonlinear_adv(or_state *S, double *fn);
/* FN of the current state (uiuj form, from om2) into fn [N][H]. */
void   or_nonlinear_uiuj(or_state *S, double *fn);

/* Raw array access (copies). */
void or_get_om2(const or_state *S, double *out);
void or_set_om2(or_state *S, const double *in);
void or_get_fnm1(const or_state *S, double *out);
void or_set_fnm1(or_state *S, const double *in);
void or_get_ur(const or_state *S, int comp, double *out);
void or_get_uc(const or_state *S, int comp, double *out);
void or_set_uc(or_state *S, int comp, const double *in);

void or_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
