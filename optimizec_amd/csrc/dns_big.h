// dns_big.h -- launch interface of the multi-pass time step (dns_big.hip) for grids
// whose lines do not fit the fused LDS-resident kernels.  Internal to the library.
//
// Layouts (elements of the transform precision unless noted; hs, ms: padded row strides
// of the H-wide and M-wide rows):
//   om, fn  [kx/8][iz][kx%8] FP64, column tiles of 8 modes (dns_state_pos; N x 8 tiles
//           elements, the last tile padded): a z pass touches one tile's N x 128 bytes
//   Y       [f][kx/BZ][z][kx%BZ], 2 or 4 fields of M x hs: column tiles (z passes only)
//   Z       [f][z][kx], 2 or 4 fields, M rows of stride hs
//   X       [p][z][x1], 1 or 2 planes, M rows of stride ms
//   Yq, Qh  [q][.], M/2 rows of stride ms
//   W       [kx/BZ][z][kx%BZ], M x hs: column tiles (z passes only)
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

#include "dns_kernels.h"
#include "dns_types.h"

enum {
    BIG_ZINV_A = 0,    // om -> Y (fields.pairs pairs of fields)                 b = Y
    BIG_ZINV_B = 1,    // Y -> Z                                                 a = Y, b = Z
    BIG_XINV_A = 2,    // Z -> X (planes 1: u + i w; 2: also w_x + i w_z)       a = Z, b = X
    BIG_XINV_B = 3,    // X -> Yq (product, first forward step)                  a = X, b = Yq
    BIG_XINV_CFL = 4,  // X (plane 0) -> cflrow[workgroup] maxima               a = X
    BIG_XFWD_B = 5,    // Yq -> Qh                                               a = Yq, b = Qh
    BIG_ZFWD_A = 6,    // Qh -> W                                                a = Qh, b = W
    BIG_ZFWD_B = 7,    // W -> FN, CN/AB2 on om, fn                              a = W
    BIG_LINE_INV = 8,  // FP64: `lines` natural-order inverse line FFTs a -> b (scratch)
    BIG_ROWS = 9,      // FP64: Z fields 0 (+ i 1 when planes == 2) -> rows R [z][x]  a = Z, b = R
    BIG_ROWS_PHYS = 10,  // FP64: rows a -> out0 (Re, or |.| with out_mode 1)
    BIG_ENERGY = 11,   // om -> per-tile E, Z partials out0[tiles], out1[tiles]
    BIG_STATE_TILE = 12,   // host-layout rows a [iz][kx] (N x H) -> state tiles b
    BIG_STATE_UNTILE = 13  // state tiles a -> host-layout rows b
};

// the state's column tiles: element (iz, kx) of om, fn on the multi-pass path
__host__ __device__ inline size_t dns_state_pos(int iz, int kx, int N)
{
    return (size_t)(kx >> 3) * ((size_t)N * 8) + (size_t)iz * 8 + (kx & 7);
}
inline int dns_state_tiles(int H) { return (H + 7) / 8; }

struct BigFields {
    int f[4];   // DNS_F_* of each field (pairs (f0, f1), (f2, f3))
    int pairs;  // 1 or 2
};

struct BigCnArgs {
    double divxz;   // 1 / M^2
    double VT;      // visc * dt / 2
    double c2, c3;  // dt (2 + cnm1) / 2, -dt cnm1 / 2
};

struct BigReq {
    int what = 0;
    int fp32 = 0;   // element type of the intermediates and tw: cf (1) or cd (0)
    int engine = 0; // with fp32: line-transform stages 0 vector (cf), 1 BF16x3 (cm3), 2 FP16 (cmh)
    int H = 0;
    hipStream_t stream = nullptr;
    const void *tw = nullptr;  // W_M^t, t in [0, M)
    dnsfft::cd *om = nullptr, *fn = nullptr;
    const void *a = nullptr;
    void *b = nullptr;
    void *scratch = nullptr;
    size_t lines = 0;
    size_t hs = 0, ms = 0;
    int planes = 2;
    // FP64 step passes 4, 5, 7 in the u_i u_j form: pass 4 reads X plane 0 (u + i w) and writes
    // Yq planes D = u^2 - w^2, C = u w (M / 2 rows each); pass 5 takes both (M rows); pass 6
    // runs once per plane; pass 7 reads W fields D, C (M hs apart) and forms FN from both
    int uiuj = 0;
    BigFields fields{};
    BigCnArgs cn{};
    double *cflrow = nullptr;
    double DX = 0.0, DZ = 0.0;
    double *out0 = nullptr, *out1 = nullptr;
    int out_mode = 0;
};

// 1 if line length M = 3N/2 has a built split
int dns_big_supported_m(int M);
// 1 if the matrix-core line transforms (engine 1, 2) are built for line length M
int dns_big_mfma_supported_m(int M);
// partial CFL maxima written by BIG_XINV_CFL (one per workgroup: two row pairs x C4 k1 lines)
int dns_big_cfl_parts(int M);
hipError_t dns_big_launch(const BigReq &r);
