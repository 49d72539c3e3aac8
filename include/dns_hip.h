/*
 * dns_hip.h -- C ABI of the MI355X (gfx950) 2D pseudo-spectral vorticity solver.
 *
 * This is the boundary the host driver (mem_hip), the Python mirror
 * (optimizec_amd) and any FFI caller bind.  Plain C types only: no HIP, torch or
 * C++ types cross it.  Every call returns a dns_status; the text of the last error
 * of a handle is available from dns_last_error().  A handle is single-threaded (no
 * internal locking); all device work of a handle is queued on one HIP stream and
 * calls that hand results to the host synchronize that stream.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * mannetroll/optimizec, directory cuda/).
 *
 * Data layouts on the host side of the ABI follow the reference:
 *   om2 / fnm1 : [iz][ix], N rows x N/2 complex (interleaved re, im doubles),
 *                kz = iz for iz < N/2, iz - N otherwise; kx = ix   (cuda_dns.h:75-76)
 *   physical planes : [z][x], M x M doubles, M = 3N/2            (cuda_dns.h:33-34)
 */
#ifndef DNS_HIP_H
#define DNS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dns_handle dns_handle;

typedef enum {
    DNS_OK = 0,
    DNS_E_ARG = 1,          /* invalid argument or size */
    DNS_E_UNSUPPORTED = 2,  /* grid size or option not built into this library */
    DNS_E_HIP = 3,          /* HIP runtime error (message in dns_last_error) */
    DNS_E_NOMEM = 4,        /* device or host allocation failed */
    DNS_E_STATE = 5,        /* call out of order (e.g. step before an initial condition) */
    DNS_E_IO = 6,           /* file could not be opened / read / written */
    DNS_E_FORMAT = 7        /* restart file rejected (magic, version, sizes, dims, counts) */
} dns_status;

typedef enum {
    DNS_IC_PAO_KM3 = 0,  /* the reference PAO-k^-3 spectrum (pao_km3_cuda.cu:59-498) */
    DNS_IC_K4EXP = 1     /* seeded random phases, E(k) ~ (k/K0)^4 exp(-2 (k/K0)^2), unit energy */
} dns_ic_kind;

typedef enum {
    DNS_PREC_FP64 = 0,  /* everything in FP64                                              */
    DNS_PREC_FP32 = 1,  /* transforms and their intermediates in FP32; state om2/fnm1, the
                           CN/AB2 update, CFL/dt, energy/enstrophy and twiddle generation
                           in FP64 (N >= 128)                                                */
    DNS_PREC_BF16X3 = 2,/* as DNS_PREC_FP32, with every radix stage (radix <= 16) of the line
                           transforms run as a DFT-matrix product on the matrix cores: three
                           BF16 products (hi*hi + hi*lo + lo*hi), FP32 accumulation
                           (N = 256, 512, 1024, 2048, 4096; N = 29400 on the multi-pass path) */
    DNS_PREC_FP16 = 3,  /* as DNS_PREC_BF16X3 with one FP16 product per stage, each butterfly's
                           inputs scaled by a power of two into (-1, 1) (same sizes)          */
    DNS_PREC_FP16_STORE = 4 /* as DNS_PREC_FP32 (vector FP32 stages), with the intermediates
                           between the step's two kernels (the z-inverted fields and the
                           x-forward product) stored as FP16 pairs under a power-of-two scale
                           per stored line: half the bytes of DNS_PREC_FP32 (N = 4096, fused
                           path)                                                              */
} dns_precision;

typedef enum {
    DNS_PATH_AUTO = 0,       /* fused kernels when built for H = N/2, else the multi-pass path   */
    DNS_PATH_FUSED = 1,      /* LDS-resident row and column kernels (N <= 4096)                  */
    DNS_PATH_MULTIPASS = 2,  /* batched four-step line FFTs through HBM (large N, e.g. 29400)    */
    DNS_PATH_FUSED_UIUJ = 3, /* the fused kernels with the nonlinear term in the reference's u_i u_j
                                (divergence) form, FN from (u^2 - w^2)^ and (u w)^
                                (step2b_cuda.cu:137-173, step3_cuda.cu:75-81) instead of the
                                advective u . grad(omega): half the z-inverted fields, two products
                                instead of one; FP64 at N = 4096                                    */
    DNS_PATH_MULTIPASS_UIUJ = 4 /* the multi-pass path in the u_i u_j form (FP64)                  */
} dns_path;

typedef enum {
    DNS_FIELD_U = 0,        /* physical u (UR_full comp 0)                     */
    DNS_FIELD_W = 1,        /* physical w (UR_full comp 1)                     */
    DNS_FIELD_OMEGA = 2,    /* physical vorticity (OM2PHYS, om2phys_cuda.cu:153) */
    DNS_FIELD_PSI = 3,      /* stream function (dnsCudaStreamFunc, om2phys_cuda.cu:258) */
    DNS_FIELD_KINETIC = 4   /* |u| (dnsCudaKinetic, om2phys_cuda.cu:358)        */
} dns_field;

/* Configuration of dnsCudaCreate(S, N, Re, K0) (cuda_dns.h:108) plus the new knobs. */
typedef struct {
    int    N;          /* base grid, multiple of 4; M = 3N/2 physical points per direction */
    double Re;         /* Reynolds number, visc = 1/Re (pao_km3_cuda.cu:380-382)            */
    double K0;         /* spectrum peak wavenumber                                          */
    double cflnum;     /* CFL number (mem_cuda.cu:1755, S.cflnum = CFL)                     */
    int    precision;  /* dns_precision                                                     */
    int    device;     /* HIP device ordinal                                                */
    int    path;       /* a dns_path value, 0 = auto                                         */
} dns_config;

/* The scalar part of DnsDeviceState (cuda_dns.h:39-69), in FP64. */
typedef struct {
    int    N, M, H, NK;
    int    it, ifn;
    double Re, K0, visc, t, dt, cn, cnm1, cflnum, cflm;
} dns_scalars;

/* ---- life cycle ------------------------------------------------------ */
/* Replaces dnsCudaCreate (cuda_dns.h:108; dns_cuda.cu:204-391): allocates the
 * state and every work buffer on the device.  Scalars start as mem_cuda.cu:1760-1765
 * (it=0, dt=0, cn=1, cnm1=0, ifn=1). */
dns_status dns_create(const dns_config *cfg, dns_handle **out);
/* Replaces dnsCudaDestroy (cuda_dns.h:109; dns_cuda.cu:393-423). */
void dns_destroy(dns_handle *h);
/* Text of the last error on h (or of the last failed dns_create when h is NULL). */
const char *dns_last_error(const dns_handle *h);
/* 1 if this library was built with kernels for base grid N (either path), else 0. */
int dns_supported_n(int N);
/* 1 if the given dns_path value is built for base grid N, else 0. */
int dns_supported_path(int N, int path);

/* ---- initial condition ------------------------------------------------ */
/* Replaces dnsCudaInit = PAO + CALCOM (cuda_dns.h:123; dns_cuda.cu:591-607): sets
 * visc, om2 (and fnm1 = 0), and prepares the inverse transforms of om2.  For
 * DNS_IC_PAO_KM3 the printed PAO block (pao_km3_cuda.cu:405-424) is written into
 * log (may be NULL).  seed is used by DNS_IC_K4EXP only. */
dns_status dns_init_ic(dns_handle *h, int kind, uint64_t seed, char *log, int logcap);

/* ---- time stepping ------------------------------------------------------ */
/* Initial NEXTDT after the first STEP2A (mem_cuda.cu:1784-1788; dns_cuda.cu:562-570):
 * dt = cflnum / CFLM / pi.  Requires it == 0 and dt == 0 (returns DNS_E_STATE otherwise). */
dns_status dns_start(dns_handle *h);
/* nsteps iterations of the mem_cuda loop body (mem_cuda.cu:1848-1857): it += 1,
 * STEP2B -> STEP3 -> STEP2A (here: row pass, then fused column pass), t += dt_old,
 * cnm1 = cn.  Does not call NEXTDT; asynchronous on the handle's stream. */
dns_status dns_step(dns_handle *h, int nsteps);
/* Replaces next_dt_gpu (dns_cuda.cu:553-580): CFLM = max |u|/DX + |w|/DZ over the
 * M x M grid of the current state; it==0 && dt==0: dt = cflnum/CFLM/pi; otherwise
 * CFL = CFLM dt pi, cn = 0.8 + 0.2 cflnum/CFL, dt *= cn.  Synchronizes. */
dns_status dns_next_dt(dns_handle *h);

/* ---- diagnostics -------------------------------------------------------- */
/* E = sum m |om2|^2 / K^2, Z = sum m |om2|^2 (m = 1 for kx = 0, else 2); the
 * reference energy convention Q2 (pao_km3_cuda.cu:366-378).  Synchronizes. */
dns_status dns_energy_enstrophy(dns_handle *h, double *E, double *Z);
/* computePalOverEnsKmax2 (cuda/mem_cuda.cu:872-935), the input of the driver's ADAPT_VISC
 * controller: sum w K^2 |om|^2 / ((sum w |om|^2 - |om(0,0)|^2) max K^2) with w = 1 on the
 * kx = 0 and kx = N/2 - 1 columns and 2 elsewhere; 0 when the enstrophy sum is not
 * positive.  Reduced on the device (the reference copies om2 to the host). */
dns_status dns_pal_over_ens_kmax2(dns_handle *h, double *out);

/* OM2PHYS + computeDisplaySigma (mem_cuda.cu:820-870): std-dev of the 8-bit
 * quantized physical vorticity.  Synchronizes. */
dns_status dns_display_sigma(dns_handle *h, double *sigma);
/* Physical plane [z][x] (M x M doubles) of a dns_field into host memory. */
dns_status dns_fetch_field(dns_handle *h, int field, double *host);

/* ---- state access -------------------------------------------------------- */
dns_status dns_get_scalars(const dns_handle *h, dns_scalars *out);
/* Sets it, ifn, Re, K0, visc, t, dt, cn, cnm1, cflnum, cflm (N, M, H, NK ignored). */
dns_status dns_set_scalars(dns_handle *h, const dns_scalars *in);
dns_status dns_get_om2(dns_handle *h, double *host);
dns_status dns_set_om2(dns_handle *h, const double *host);
dns_status dns_get_fnm1(dns_handle *h, double *host);
dns_status dns_set_fnm1(dns_handle *h, const double *host);

/* ---- restart.bin (mem_cuda.cu:1019-1394) ---------------------------------- */
/* Writes RestartHeader v2 with real_bytes=8, cplx_bytes=16, then UR_full (u, w,
 * omega physical, 3 x M x M), om2 and fnm1. */
dns_status dns_save_restart(dns_handle *h, const char *path);
/* Reads v1 or v2 files with 4- or 8-byte reals (FP32 files are upcast); rejects bad
 * magic, version, header size, real/cplx size, dims or counts with DNS_E_FORMAT. */
dns_status dns_load_restart(dns_handle *h, const char *path);

/* ---- host-only helpers (no device needed) ---------------------------------- */
/* PAO-k^-3 spectrum as om2 [N][N/2] complex (after CALCOM); returns the final LCG
 * seed (the printed "PAO seed"), or -1 on bad arguments. */
int dns_ic_pao_host(int N, double Re, double K0, double *om2, char *log, int logcap, double *visc);
/* k^4 exp initial condition as om2 [N][N/2] complex. */
dns_status dns_ic_k4exp_host(int N, double K0, uint64_t seed, double *om2);

/* ---- measurement ---------------------------------------------------------- */
/* HIP-event timing of hot-kernel launches on the handle's stream: enabled <= 0 turns it
 * off, k >= 1 times the 1st, (k+1)-th, (2k+1)-th ... launch of each kernel (1 = every
 * launch); a pair of events costs a few microseconds of stream time. */
dns_status dns_set_kernel_timing(dns_handle *h, int enabled);
/* Per-kernel totals since the last reset: names[i] (static strings), launches,
 * total milliseconds, algorithmic bytes per launch.  *n is capacity in, count out. */
dns_status dns_kernel_timing(dns_handle *h, const char **names, int64_t *launches,
                             double *total_ms, double *bytes_per_launch, int *n);
dns_status dns_reset_kernel_timing(dns_handle *h);
/* ---- allocation-free step --------------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP stream capture, tight serving loops, harnesses that forbid hipMalloc):
 * nsteps iterations of the dns_step loop body (FP64 fused kernels, bitwise the results of
 * dns_step on a handle holding the same state and scalars) on caller-owned device arrays
 * om2, fnm1 in the reference's device layout [iz][ix] (N x N/2 complex, interleaved
 * doubles; cuda_dns.h:75-76), enqueued on `stream` (a hipStream_t; NULL = the null
 * stream) of the current device.  Reads N, visc, dt, cn, cnm1, t, it of *sc and updates
 * it, t, cnm1 as dns_step does; dt must already be set (dns_start / dns_next_dt on a
 * handle, or the caller's own NEXTDT).  N must be one of 16, 32, 64, 128, 256, 384,
 * 512 (DNS_NOALLOC_MAX_N): larger N returns DNS_E_ARG.  All work buffers live in one
 * file-scope device array of DNS_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends and shared by every call on that GPU: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialized by the caller.  The first call
 * on a device does a one-time set-up (symbol lookup and a 33 KB synchronous twiddle-table
 * copy): make it outside any stream capture.  No hipMalloc, hipFree or allocating call
 * is made.  Errors: dns_last_error(NULL). */
#define DNS_NOALLOC_MAX_N 512
#define DNS_NOALLOC_SCRATCH_BYTES (20u * 1024u * 1024u)
dns_status dns_step_noalloc(dns_scalars *sc, double *om2, double *fnm1, int nsteps, void *stream);

/* Blocks until all queued work of the handle is done. */
dns_status dns_synchronize(dns_handle *h);
/* Device bytes allocated by the handle. */
size_t dns_device_bytes(const dns_handle *h);

#ifdef __cplusplus
}
#endif
#endif /* DNS_HIP_H */
