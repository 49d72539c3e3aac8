// cnn_kernels.h -- launch interface of the closure network's kernels (cnn_kernels.hip).
// Internal to the library.
//
// Activations are channels-last ("NHWC") BF16, [pixel][channel] with the pixel index
// p = (b n + y) n + x; pre-activations, outputs and gradients of the weights are FP32.
// The kernels are implicit GEMMs on the gfx950 bf16 matrix instruction
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

#include "dns_types.h"

// epilogues of the convolution kernel
enum { CONV_BIAS_GELU = 0, CONV_BIAS = 1, CONV_GELU_GRAD = 2 };

// out = conv3x3_circular(in, W) (+ epilogue).  in: [P][cinp] bf16; W: [coutp][kp] bf16 with
// k = tap * cinp + ci, tap = 3 ky + kx, kp = 9 cinp rounded up to 32 (zero padded).
//   CONV_BIAS_GELU: z = acc + bias -> outz [P][coutp] (FP32), outa = GELU(z) (bf16)
//   CONV_BIAS:      outz[p * cout_real + co] = acc + bias for co < cout_real
//   CONV_GELU_GRAD: outa = acc * GELU'(zprev) (bf16)  (data gradient of the next layer)
hipError_t cnn_launch_conv(int cinp, int coutp, int epi, const uint16_t *in, const uint16_t *W, const float *bias,
                           int n, int P, float *outz, uint16_t *outa, const float *zprev, int cout_real,
                           hipStream_t s);
// partial weight/bias gradients of one layer over pixel blocks of `ppb` pixels:
// part[blk][coutp][9 cinp] = sum_p dz[p][co] in[p + off(tap)][ci], partb[blk][coutp]
hipError_t cnn_launch_wgrad(int cinp, int coutp, const uint16_t *dz, const uint16_t *in, int n, int P, int ppb,
                            float *part, float *partb, hipStream_t s);
// sum of the partials into the flat gradient (layout of the parameter vector) of a
// layer with cin x cout real channels at offset off
hipError_t cnn_launch_wgrad_reduce(int cinp, int coutp, int cin, int cout, const float *part, const float *partb,
                                   int blocks, float *grad, size_t off, hipStream_t s);
// y, t: [P]; per-block sums of (y - t)^2 into lossp (double); dz[p][0..15] = bf16(2 (y - t) / P), 0
hipError_t cnn_launch_mse(const float *y, const float *t, int P, double *lossp, int blocks, uint16_t *dz,
                          hipStream_t s);
// Adam on the FP32 master vector: c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
hipError_t cnn_launch_adam(float *p, const float *g, float *m, float *v, int count, float lr, float b1, float b2,
                           float eps, float c1, float c2, hipStream_t s);
// master layer weights (at offset off) -> forward [coutp][kp] and data-gradient [cinp_b][kpb]
// bf16 layouts (Wb may be null), biases -> bias [coutp] FP32
hipError_t cnn_launch_pack_layer(const float *p, size_t off, int cin, int cout, int cinp, int coutp, uint16_t *Wf,
                                 uint16_t *Wb, float *bias, hipStream_t s);
// x: [B][3][n][n] FP32 (channel-major) -> a0 [P][8] bf16 (channels 3..7 zero)
hipError_t cnn_launch_pack_input(const float *x, int B, int n, uint16_t *a0, hipStream_t s);
// training minibatch from stored pairs ([slot][member][4][n][n] FP32): sample s of
// iteration `iter` is pair (slot, member) = hash(seed, iter, s) mod (stored x count);
// per-sample scaling of closure_scaling.hpp (small-scale fields over their vorticity rms)
hipError_t cnn_launch_gather_pairs(const float *pairs, int count, int stored, int n, int B, uint64_t seed, int iter,
                                   uint16_t *a0, float *t, hipStream_t s);
