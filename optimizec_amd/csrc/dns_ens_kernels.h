// dns_ens_kernels.h -- launch interface of the ensemble's device-side scalar logic and
// training-pair kernels (dns_ens_kernels.hip).  Internal to the library.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

#include "dns_types.h"

// per-member scalars kept on the device (the dns_scalars fields the loop changes)
struct EnsScalars {
    int it, ifn;
    double t, dt, dt_old, cn, cnm1, visc, cflnum, cflm;
};

// it += 1, dt_old = dt, CN/AB2 coefficients (VT, c2, c3) of the step into coef[3 b ..]
hipError_t ens_launch_pre(EnsScalars *sc, double *coef, int count, hipStream_t s);
// t += dt_old, cnm1 = cn (after the step)
hipError_t ens_launch_post(EnsScalars *sc, int count, hipStream_t s);
// CFLM = max of the member's M row maxima, then the NEXTDT logic (dns_cuda.cu:553-580)
hipError_t ens_launch_nextdt(EnsScalars *sc, const double *cflrow, int M, int count, hipStream_t s);
// dst[i] = src[i], n complex values
hipError_t ens_launch_copy(const dnsfft::cd *src, dnsfft::cd *dst, size_t n, hipStream_t s);
// coarse spectra of count members: om_c = F(om_f) on the n-grid layout [kx][iz], fn_c = 0
hipError_t ens_launch_truncate(const dnsfft::cd *om_f, int N, dnsfft::cd *om_c, dnsfft::cd *fn_c, int n, int count,
                               hipStream_t s);
// the 4 channels of one training pair per member into pairs + (slot * count + b) * 4 n^2,
// slot read from *slot (no write when *slot >= max_pairs)
hipError_t ens_launch_pairs(const dnsfft::cd *om_c, const dnsfft::cd *fn_f, int N, const dnsfft::cd *fn_c, int n,
                            int count, float *pairs, const int *slot, int max_pairs, hipStream_t s);
// *slot += 1
hipError_t ens_launch_bump(int *slot, hipStream_t s);
