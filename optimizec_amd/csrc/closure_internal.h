// closure_internal.h -- library-internal interface between the closure network
// (dns_closure.cpp), the grid transforms of the closure forcing (closure_kernels.hip)
// and the solver (dns_api.cpp).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

#include "../../include/dns_hip.h"
#include "dns_types.h"

struct dns_cnn;

// forward pass of B samples of n x n from the network's input buffer (filled by the
// caller) into its output buffer, on stream s
dns_status cnn_forward_device(dns_cnn *c, int B, int n, hipStream_t s);
uint16_t *cnn_input_buffer(dns_cnn *c);  // [P][8] bf16
float *cnn_output_buffer(dns_cnn *c);    // [P] FP32
dns_status cnn_reserve(dns_cnn *c, size_t P);
int cnn_device(dns_cnn *c);
const char *cnn_error(dns_cnn *c);

// closure forcing transforms on the N x N grid (N in 32 .. 1024, powers of two)
int closure_grid_supported(int N);
// om [kx][iz] (H x N) -> fields omega, u, w on the grid [3][N][N] FP32; az: work array
// [3][N][H] complex
hipError_t closure_launch_to_grid(const dnsfft::cd *om, int N, const dnsfft::cd *twN, dnsfft::cd *az, float *fields,
                                  hipStream_t s);
// small-scale fields hp [3][N][N] (closure_scaling.hpp), their vorticity rms w_s,
// scale = (1 / w_s, 1 / (w_s D), w_s^2), network input a0 [N^2][8] bf16 = (omega' scale0,
// u' scale1, w' scale1, 0 ...); part: closure_pack_blocks() doubles
int closure_pack_blocks();
hipError_t closure_launch_pack(const float *fields, int N, float *hp, double *part, double *scale, uint16_t *a0,
                               hipStream_t s);
// shaping of the forcing: band ramp (0 for |k| <= k_lo, 1 for |k| >= k_hi) and, with om
// set, the dissipative limiter f = -nu om with 0 <= nu dt <= nu_dt_max (nu from the
// projection of the network's forcing on -om); the mean and the kz = -N/2 row are 0
struct ClosureShape {
    double k_lo = 0.0, k_hi = 0.0;
    const dnsfft::cd *om = nullptr;
    double dt = 0.0, nu_dt_max = 0.0;
};
// network output y [N][N] (x scale[2]) -> spectral forcing fext [kx][iz] (H x N);
// bz: work array [N][H] complex
hipError_t closure_launch_from_grid(const float *y, const double *scale, int N, const dnsfft::cd *twN,
                                    dnsfft::cd *bz, dnsfft::cd *fext, const ClosureShape &sh, hipStream_t s);
