/*
 * dns_closure.h -- C ABI of the learned subgrid closure: a small convolutional network
 * trained on the GPU from the ensemble's filtered-DNS pairs and applied inside a coarse
 * run as an extra forcing of the vorticity equation.
 *
 * There is no counterpart in the reference (mannetroll/optimizec has no closure and no
 * learned component).  Network: four 3x3 convolutions with circular padding,
 * 3 -> 32 -> 32 -> 32 -> 1 channels, GELU (erf form) after the first three, 19 681
 * parameters; FP32 master weights, BF16 matrix-core compute with FP32 accumulation,
 * MSE loss, Adam (beta1 0.9, beta2 0.999, eps 1e-8).
 *
 * Parameter vector: for each layer W[co][ci][ky][kx] then b[co] (FP32); the layer
 * computes out[y][x] = b + sum W[ky][kx] in[(y+ky-1) mod n][(x+kx-1) mod n].
 * Network input x: [B][3][n][n] FP32, output / target: [B][n][n] FP32, in the network's
 * units: with f' = f minus its mean over the 5 x 5 periodic box of grid points around the
 * point (the small-scale part of a field), w_s = rms of omega' and D = 2 pi / n the grid
 * spacing, channels (omega' / w_s, u' / (w_s D), w' / (w_s D)) and subgrid term
 * tau / w_s^2.  The subgrid term of near-cutoff interactions scales as w_s^2, so a grid
 * whose cutoff carries little vorticity receives a small forcing.
 */
#ifndef DNS_CLOSURE_H
#define DNS_CLOSURE_H

#include <stddef.h>
#include <stdint.h>

#include "dns_ensemble.h"
#include "dns_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dns_cnn dns_cnn;

/* Parameters: uniform in +-sqrt(3 / fan_in) from a splitmix64 stream of seed; biases 0. */
dns_status dns_cnn_create(int device, uint64_t seed, dns_cnn **out);
void dns_cnn_destroy(dns_cnn *c);
const char *dns_cnn_last_error(const dns_cnn *c);
int dns_cnn_param_count(void);
dns_status dns_cnn_get_params(dns_cnn *c, float *host);
/* Sets the master weights and clears the Adam moments and step count. */
dns_status dns_cnn_set_params(dns_cnn *c, const float *host);
/* y = net(x) for B samples of n x n (n a multiple of 4, B n^2 a multiple of 32). */
dns_status dns_cnn_forward(dns_cnn *c, const float *x, int B, int n, float *y);
/* MSE loss of net(x) against t and its gradient (flat parameter layout), no update. */
dns_status dns_cnn_loss_grad(dns_cnn *c, const float *x, const float *t, int B, int n, double *loss, float *grad);
/* One Adam step on (x, t); *loss (may be NULL) is the loss before the step. */
dns_status dns_cnn_adam_step(dns_cnn *c, const float *x, const float *t, int B, int n, float lr, double *loss);
/* iterations Adam steps on minibatches of `batch` pairs drawn (seeded hash) from the
 * pairs an ensemble holds; for every 4 iterations, groups are visited in proportion to
 * their member counts.  loss_history (may be NULL) receives one loss per iteration. */
dns_status dns_cnn_train_on_ensemble(dns_cnn *c, dns_ens *e, int iterations, int batch, float lr, uint64_t seed,
                                     double *loss_history);

/* Coarse runs: with a network attached, every dns_step adds the network's subgrid
 * forcing of the current state (fields on the N x N grid, per-step velocity scale) to
 * the nonlinear term before the CN/AB2 update.  FP64 fused path only; NULL detaches. */
dns_status dns_set_closure(dns_handle *h, dns_cnn *c);
/* Wavenumber band of the forcing: each mode is weighted by 0 for |k| <= k_lo, 1 for
 * |k| >= k_hi, linearly between (default 0, 0: all modes); the mean vorticity mode and
 * the kz = -N/2 row never receive forcing (the subgrid term is a divergence). */
dns_status dns_set_closure_band(dns_handle *h, double k_lo, double k_hi);
/* Dissipative limiter (default on, nu_dt_max 0.5): each mode's forcing is replaced by
 * -nu omega-hat, nu = the damping rate of the network's forcing projected on -omega-hat,
 * clipped to 0 <= nu dt <= nu_dt_max -- no backscatter and no explicit damping beyond
 * the AB2 step's stable range, so a network outside its training range degrades to the
 * plain run instead of destabilizing it.  enabled = 0 applies the raw forcing. */
dns_status dns_set_closure_limiter(dns_handle *h, int enabled, double nu_dt_max);
/* The forcing the attached network gives for the current state, [iz][ix] like om2. */
dns_status dns_closure_forcing(dns_handle *h, double *host);

#ifdef __cplusplus
}
#endif
#endif /* DNS_CLOSURE_H */
