/*
 * dns_ensemble.h -- C ABI of the batched ensemble: many independent runs of the 2D
 * vorticity DNS (uneven N, seed, K0, Re) advanced together on one GPU, with filtered-DNS
 * training pairs for a subgrid closure emitted every few steps.
 *
 * There is no counterpart in the reference (mannetroll/optimizec runs one simulation
 * per process, cuda/mem_cuda.cu:1654-2082).  Each member follows exactly the loop of
 * the single-run API (include/dns_hip.h: dns_step, then dns_next_dt when it == 1 or
 * it % pair_every == 0); its state is bitwise equal to a single run with the same
 * (N, Re, K0, cflnum, seed).  Members of one N share every launch (batched kernels);
 * the device keeps all scalars (it, t, dt, cn, cnm1, CFLM), so a block of pair_every
 * steps needs no host synchronization and is captured once into a HIP graph and
 * replayed.
 *
 * Training pair e of a member (emitted from the state at it = e * pair_every, before
 * the step that follows it), on the coarse grid n x n with n = N / filter_ratio and
 * the sharp spectral filter F that keeps |kx|, |kz| < n/2:
 *   channel 0  filtered vorticity  F(omega)
 *   channel 1  filtered u          F(u)
 *   channel 2  filtered w          F(w)
 *   channel 3  subgrid term        tau = F(-(u . grad omega)) - NL_n(F(omega))
 * where NL_n is the dealiased nonlinear term of the n-grid solver (the term a coarse
 * run computes itself), so a coarse run that adds tau reproduces the filtered DNS
 * tendency.  Values are FP32, channel-major [4][z][x] with x fastest, on the points
 * x_j = 2 pi j / n of the coarse grid.
 */
#ifndef DNS_ENSEMBLE_H
#define DNS_ENSEMBLE_H

#include <stddef.h>
#include <stdint.h>

#include "dns_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dns_ens dns_ens;

/* One member: base grid N (fused-kernel sizes), Re, K0, CFL number and IC seed
 * (DNS_IC_K4EXP initial condition). */
typedef struct {
    int      N;
    double   Re;
    double   K0;
    double   cflnum;
    uint64_t seed;
} dns_ens_member;

typedef struct {
    int                   nmembers;
    const dns_ens_member *members;
    int                   pair_every;    /* steps between NEXTDT calls and training pairs (e.g. 20) */
    int                   filter_ratio;  /* coarse grid n = N / filter_ratio (e.g. 8)             */
    int                   max_pairs;     /* training pairs kept per member (device memory)        */
    int                   use_graph;     /* 1: capture a block of pair_every steps in a HIP graph
                                            and replay it; 2: capture each dns_ens_advance call
                                            whole (all its blocks, the first block's NEXTDT after
                                            step 1 included) into one graph, launched once and
                                            reused by later calls of the same length; 0: none */
    int                   device;
    int                   autotune;      /* 1: dns_ens_init times the full step (row pass +
                                            column kernel) of each launch configuration whose
                                            results are bitwise the default's and keeps the
                                            fastest; 2: also the radix-16-plan configurations
                                            (results then agree with single runs to roundoff);
                                            3 + v: every size runs configuration v (0 <= v < 9
                                            where the size has nine, else 0), untimed; 0:
                                            configuration 0                                     */
} dns_ens_config;

/* Allocates every member's state and work arrays on the device; no IC yet. */
dns_status dns_ens_create(const dns_ens_config *cfg, dns_ens **out);
void dns_ens_destroy(dns_ens *e);
/* Text of the last error on e (or of the last failed dns_ens_create when e is NULL). */
const char *dns_ens_last_error(const dns_ens *e);
/* Each member's IC (k^4 exp spectrum, its seed), then the initial NEXTDT (dt from CFLM). */
dns_status dns_ens_init(dns_ens *e);
/* Replaces member m's om2 (host layout of dns_set_om2) and zeroes its FNM1; the member's
 * scalars are left as they are (set them with dns_ens_set_scalars). */
dns_status dns_ens_set_om2(dns_ens *e, int m, const double *host);
/* Advances every member by nsteps (a multiple of pair_every), emitting the training pairs
 * of the states it = k * pair_every passed on the way; asynchronous. */
dns_status dns_ens_advance(dns_ens *e, int nsteps);
dns_status dns_ens_synchronize(dns_ens *e);
dns_status dns_ens_get_scalars(dns_ens *e, int m, dns_scalars *out);
dns_status dns_ens_set_scalars(dns_ens *e, int m, const dns_scalars *in);
dns_status dns_ens_get_om2(dns_ens *e, int m, double *host);
dns_status dns_ens_get_fnm1(dns_ens *e, int m, double *host);
/* Number of training pairs emitted so far (the same for every member), and member m's
 * coarse grid size n. */
dns_status dns_ens_pair_info(dns_ens *e, int m, int *count, int *n);
/* Pairs first .. first + count - 1 of member m into host: count x 4 x n x n floats. */
dns_status dns_ens_get_pairs(dns_ens *e, int m, int first, int count, float *host);
/* Launch configuration chosen for member m's size (0 = default) and the time of one
 * batched step of each built configuration (ms[i], i < *n; 0 where not timed).  The
 * configurations (column tile x column threads, row threads, opaque FFT index, rows per
 * row workgroup, radix plan) are listed in DESIGN.md; all but the radix-16 plans (7, 8)
 * give bitwise identical results. */
dns_status dns_ens_tuning(dns_ens *e, int m, int *variant, double *ms, int *n);
/* Device bytes held by the ensemble. */
size_t dns_ens_device_bytes(const dns_ens *e);

#ifdef __cplusplus
}
#endif
#endif /* DNS_ENSEMBLE_H */
