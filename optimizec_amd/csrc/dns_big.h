// dns_big.h -- launch interface of the multi-pass time step (dns_big.hip) for grids
// whose lines do not fit the fused LDS-resident kernels.  Internal to the library.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

#include "dns_kernels.h"
#include "dns_types.h"

enum {
    BIG_SPEC = 0,     // om -> nfields padded spectra as z-lines (a: out)
    BIG_FFT_INV = 1,  // `lines` inverse line FFTs of length M: a -> b (scratch: pass-A output)
    BIG_FFT_FWD = 2,  // forward, same arguments
    BIG_PACK = 3,     // z-lines a (+ a2) -> Hermitian row spectra a + i a2 in rows b [M][M]
    BIG_PROD = 4,     // rows a = u + i w, a2 = w_x + i w_z -> row-pair products b [M/2][M]
    BIG_CFL = 5,      // rows a = u + i w -> cflrow[M]
    BIG_UNPACK = 6,   // forward row-pair spectra a -> z-lines b [H][M]
    BIG_CN = 7,       // z-forward lines a -> FN, CN/AB2 update of om, fn
    BIG_PHYS = 8      // rows a -> out0 (Re) / out1 (Im), or |a| with out_mode 1 (FP64 out)
};

struct BigSpecArgs {
    int nfields;
    int fields[4];  // DNS_F_*
};

struct BigCnArgs {
    double divxz;   // 1 / M^2
    double VT;      // visc * dt / 2
    double c2, c3;  // dt (2 + cnm1) / 2, -dt cnm1 / 2
};

struct BigReq {
    int what = 0;
    int fp32 = 0;   // element type of a, a2, b, scratch, tw: cf (1) or cd (0)
    int H = 0;
    hipStream_t stream = nullptr;
    const void *tw = nullptr;  // W_M^t, t in [0, M)
    dnsfft::cd *om = nullptr, *fn = nullptr;
    void *a = nullptr;
    const void *a2 = nullptr;
    void *b = nullptr;
    void *scratch = nullptr;
    size_t lines = 0;
    BigSpecArgs spec{};
    BigCnArgs cn{};
    double *cflrow = nullptr;
    double DX = 0.0, DZ = 0.0;
    double *out0 = nullptr, *out1 = nullptr;
    int out_mode = 0;
};

// 1 if line length M = 3N/2 has a built four-step split
int dns_big_supported_m(int M);
hipError_t dns_big_launch(const BigReq &r);
