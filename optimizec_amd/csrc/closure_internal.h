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
// om [kx][iz] (H x N) -> fields omega, u, w on the grid [3][N][N] FP32 and per-row
// sums of u^2 + w^2 (rowsum[N]); az: work array [3][N][H] complex
hipError_t closure_launch_to_grid(const dnsfft::cd *om, int N, const dnsfft::cd *twN, dnsfft::cd *az, float *fields,
                                  double *rowsum, hipStream_t s);
// scale[0] = D / U, scale[1] = 1 / U, scale[2] = U^2 / D^2 (U = rms velocity, D = 2 pi / N),
// then the network input a0 [N^2][8] bf16 = (omega scale0, u scale1, w scale1, 0 ...)
hipError_t closure_launch_pack(const float *fields, const double *rowsum, int N, double *scale, uint16_t *a0,
                               hipStream_t s);
// network output y [N][N] (x scale[2]) -> spectral forcing fext [kx][iz] (H x N), weighted
// by the band ramp (0 for |k| <= k_lo, 1 for |k| >= k_hi); the mean and the kz = -N/2 row
// are 0; bz: work array [N][H] complex
hipError_t closure_launch_from_grid(const float *y, const double *scale, int N, const dnsfft::cd *twN,
                                    dnsfft::cd *bz, dnsfft::cd *fext, double k_lo, double k_hi, hipStream_t s);
