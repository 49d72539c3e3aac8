// dns_kernels.h -- launch interface between the host API (dns_api.cpp) and the
// gfx950 kernels (dns_kernels.hip).  Internal to the library; not part of the C ABI.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

#include "dns_types.h"

// spectra the column kernel can z-invert (multipliers of omega)
enum {
    DNS_F_U = 0,      // -i gamma / K^2      (kx = 0: -i / gamma)
    DNS_F_W = 1,      //  i alfa / K^2
    DNS_F_OX = 2,     //  i alfa             (d omega / dx)
    DNS_F_OZ = 3,     //  i gamma            (d omega / dz)
    DNS_F_OMEGA = 4,  //  1
    DNS_F_PSI = 5     //  1 / K^2            (stream function)
};

struct ColArgs {
    int do_forward;     // 1: z-forward of G + CN/AB2 update before the inverse
    int nfields;        // spectra to z-invert into V (field fi at V + fi * vstride)
    int fields[4];
    size_t vstride;
    double divxz;       // 1 / M^2
    double VT;          // visc * dt / 2
    double c2, c3;      // dt (2 + cnm1) / 2, -dt cnm1 / 2
    const dnsfft::cd *fext;  // extra forcing added to FN, layout of om (single problem), or null
};

enum { LAUNCH_COL = 0, LAUNCH_ROW = 1, LAUNCH_ROW_CFL = 2, LAUNCH_ROW_PHYS = 3 };

// Independent problems of one size advanced by one launch (blockIdx.y = member b):
// member b's arrays start at the member-0 pointers plus b times these strides
// (elements), and its CN/AB2 coefficients (VT, c2, c3) are read from coef[3 b ..]
// when coef is set (the column kernel's ColArgs values otherwise).
struct BatchArgs {
    int count = 1;
    size_t s_om = 0, s_v = 0, s_g = 0, s_cfl = 0;
    const double *coef = nullptr;
};

struct LaunchReq {
    int what;
    hipStream_t stream;
    int fp32;                  // 1: transforms and V, G in FP32 (cf); 0: FP64 (cd)
    int engine;                // with fp32: 0 vector stages (cf), 1 BF16x3 (cm3), 2 FP16 (cmh) matrix-core
                               // stages, 3 vector stages with V and G stored as FP16 pairs (cs)
    // column kernel
    dnsfft::cd *om, *fn;
    const void *G;             // cd or cf
    void *V;                   // cd or cf
    const dnsfft::cd *tw;      // FP64 twiddles W_M^t
    const void *twT;           // twiddles of the transform precision
    const void *twH;           // W_H^j = twT[3 j] (j < H) stored contiguously, or null (then read
                               // from twT with stride 3)
    ColArgs col;
    // row kernels
    size_t vstride;
    void *Gw;
    double *cflrow;            // LAUNCH_ROW_CFL: per-row CFL maxima (LAUNCH_ROW: must be null)
    double DX, DZ;
    const dnsfft::cd *Va, *Vb;
    double *out0, *out1;
    int out_mode;
    BatchArgs batch;           // column and row kernels (MODE 0, 1)
    int variant;               // launch configuration (0 default; see dns_kernels_variants)
    // engine 3: the exponents of the FP16 storage -- vexp[kx] the four V fields' exponents of
    // mode kx as bytes (written by the column kernel, read by the row kernel), gexp[z] that of
    // G row z (written by the row kernel, read by the column kernel)
    unsigned *vexp;
    int *gexp;
    // the nonlinear term's form of the step's row and column kernels: 0 advective (V: u, w,
    // d_x omega, d_z omega; G: one product), 1 u_i u_j (V: u, w; G: D^ = (u^2 - w^2)^ and
    // C^ = (u w)^ at G + vstride; FP64 at H = 2048, see dns_kernels_uiuj_supported)
    int form;
};

int dns_kernels_supported(int H);
// 1 if the matrix-core line transforms (engine 1, 2) are built for H
int dns_kernels_mfma_supported(int H);
// 1 if V and G stored as FP16 pairs (engine 3) are built for H
int dns_kernels_h16_supported(int H);
// 1 if the FP64 step kernels for H are built in the u_i u_j form too (LaunchReq::form = 1)
int dns_kernels_uiuj_supported(int H);
// number of launch configurations of the row / column kernels built for H (variant
// 0 .. n-1 of LaunchReq); all give bitwise identical results
int dns_kernels_variants(int H, int fp32);
// 1 if variant v gives results bitwise those of variant 0 (the same arithmetic per element),
// 0 if it changes the radix plan (agreement to roundoff)
int dns_kernels_variant_bitwise(int v);
int dns_kernels_col_tile(int H);
hipError_t dns_kernels_launch(int H, const LaunchReq &r);
hipError_t dns_launch_energy(const dnsfft::cd *om, int N, int H, double *partE, double *partZ, hipStream_t s);
// ADAPT_VISC input partials (one per line of the state: N / 2 columns of the fused [kx][iz]
// layout, or, with tiled = 1, one per column tile of 8 modes of the multi-pass layout)
hipError_t dns_launch_palin(const dnsfft::cd *om, int N, int tiled, double *partT, double *partP, hipStream_t s);
hipError_t dns_launch_minmax(const double *f, size_t n, double *out, int blocks, hipStream_t s);
hipError_t dns_launch_pixels(const double *f, size_t n, double mn, double rng, int is_const,
                             unsigned long long *out, int blocks, hipStream_t s);
hipError_t dns_launch_transpose(const dnsfft::cd *in, dnsfft::cd *out, int rows, int cols, hipStream_t s);

// W_M^t = e^{-2 pi i t / M}, t in [0, M), in long double rounded once (host; dns_api.cpp):
// every path of the library uses this one table
void dns_host_twiddles(int M, dnsfft::cd *out);
// sets the text dns_last_error(NULL) returns (calls without a handle)
void dns_note_error(const char *msg);
