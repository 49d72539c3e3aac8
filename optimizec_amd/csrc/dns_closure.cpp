// dns_closure.cpp -- host side of the closure network (include/dns_closure.h): weights,
// Adam state, activation buffers, the forward / backward launch sequence, training on
// an ensemble's pairs.  The network has no counterpart in the reference.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dns_closure.h"
#include "closure_internal.h"
#include "cnn_kernels.h"
#include "dns_ens_internal.h"

namespace {

thread_local std::string g_cnn_create_error;

struct LayerSpec {
    int cin, cout, cinp, coutp;
};
const LayerSpec kLayers[4] = {{3, 32, 8, 32}, {32, 32, 32, 32}, {32, 32, 32, 32}, {32, 1, 32, 16}};

constexpr int kp_of(int k) { return (k + 31) / 32 * 32; }

size_t layer_offset(int l)
{
    size_t o = 0;
    for (int i = 0; i < l; ++i) o += (size_t)kLayers[i].cout * kLayers[i].cin * 9 + kLayers[i].cout;
    return o;
}

constexpr int kParams = 3 * 32 * 9 + 32 + 2 * (32 * 32 * 9 + 32) + 32 * 9 + 1;
constexpr int kMseBlocks = 1024;
constexpr int kMaxWgradBlocks = 512;

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

struct dns_cnn {
    int device = 0;
    hipStream_t stream = nullptr;
    float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
    uint16_t *Wf[4] = {}, *Wb[4] = {};
    float *bias[4] = {};
    int step = 0;
    size_t Pcap = 0;
    uint16_t *A0 = nullptr, *A1 = nullptr, *A2 = nullptr, *A3 = nullptr;
    float *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *Y = nullptr, *T = nullptr, *xstage = nullptr;
    uint16_t *dz4 = nullptr, *dz3 = nullptr, *dz2 = nullptr, *dz1 = nullptr;
    float *part = nullptr, *partb = nullptr;
    double *lossp = nullptr;
    std::vector<void *> fixed;  // allocations that live as long as the network
    std::string err;
};

namespace {

dns_status cfail(dns_cnn *c, dns_status st, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c)
        c->err = buf;
    else
        g_cnn_create_error = buf;
    return st;
}

#define CHIP(expr)                                                                                \
    do {                                                                                          \
        hipError_t x_ = (expr);                                                                   \
        if (x_ != hipSuccess) return cfail(c, DNS_E_HIP, "%s: %s", #expr, hipGetErrorString(x_)); \
    } while (0)
#define CSTEP(expr)                    \
    do {                               \
        dns_status st_ = (expr);       \
        if (st_ != DNS_OK) return st_; \
    } while (0)

template <class T>
dns_status calloc_dev(dns_cnn *c, T **p, size_t count, const char *what)
{
    void *q = nullptr;
    hipError_t x = hipMalloc(&q, sizeof(T) * (count ? count : 1));
    if (x != hipSuccess) {
        (void)hipGetLastError();
        return cfail(c, DNS_E_NOMEM, "hipMalloc(%s): %s", what, hipGetErrorString(x));
    }
    *p = static_cast<T *>(q);
    return DNS_OK;
}

void free_activations(dns_cnn *c)
{
    void *ptrs[] = {c->A0, c->A1, c->A2, c->A3, c->Z1, c->Z2, c->Z3, c->Y, c->T, c->xstage,
                    c->dz4, c->dz3, c->dz2, c->dz1, c->part, c->partb};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    c->A0 = c->A1 = c->A2 = c->A3 = nullptr;
    c->Z1 = c->Z2 = c->Z3 = c->Y = c->T = c->xstage = nullptr;
    c->dz4 = c->dz3 = c->dz2 = c->dz1 = nullptr;
    c->part = c->partb = nullptr;
    c->Pcap = 0;
}

// activation buffers for P pixels (grown, never shrunk)
dns_status ensure_capacity(dns_cnn *c, size_t P)
{
    if (P <= c->Pcap) return DNS_OK;
    // a solver with this network attached may have work in flight on its own stream
    CHIP(hipDeviceSynchronize());
    free_activations(c);
    CSTEP(calloc_dev(c, &c->A0, P * 8, "input"));
    CSTEP(calloc_dev(c, &c->A1, P * 32, "a1"));
    CSTEP(calloc_dev(c, &c->A2, P * 32, "a2"));
    CSTEP(calloc_dev(c, &c->A3, P * 32, "a3"));
    CSTEP(calloc_dev(c, &c->Z1, P * 32, "z1"));
    CSTEP(calloc_dev(c, &c->Z2, P * 32, "z2"));
    CSTEP(calloc_dev(c, &c->Z3, P * 32, "z3"));
    CSTEP(calloc_dev(c, &c->Y, P, "output"));
    CSTEP(calloc_dev(c, &c->T, P, "target"));
    CSTEP(calloc_dev(c, &c->xstage, P * 3, "input staging"));
    CSTEP(calloc_dev(c, &c->dz4, P * 16, "dz4"));
    CSTEP(calloc_dev(c, &c->dz3, P * 32, "dz3"));
    CSTEP(calloc_dev(c, &c->dz2, P * 32, "dz2"));
    CSTEP(calloc_dev(c, &c->dz1, P * 32, "dz1"));
    CSTEP(calloc_dev(c, &c->part, (size_t)kMaxWgradBlocks * 32 * 288, "weight-gradient partials"));
    CSTEP(calloc_dev(c, &c->partb, (size_t)kMaxWgradBlocks * 32, "bias-gradient partials"));
    c->Pcap = P;
    return DNS_OK;
}

dns_status check_shape(dns_cnn *c, int B, int n)
{
    if (B < 1 || n < 4 || n % 4 != 0) return cfail(c, DNS_E_ARG, "B=%d, n=%d: need B >= 1 and n a multiple of 4", B, n);
    if (((size_t)B * n * n) % 32 != 0) return cfail(c, DNS_E_ARG, "B n^2 = %zu must be a multiple of 32", (size_t)B * n * n);
    return ensure_capacity(c, (size_t)B * n * n);
}

dns_status pack_weights(dns_cnn *c)
{
    for (int l = 0; l < 4; ++l) {
        const LayerSpec &s = kLayers[l];
        CHIP(cnn_launch_pack_layer(c->p, layer_offset(l), s.cin, s.cout, s.cinp, s.coutp, c->Wf[l], c->Wb[l],
                                   c->bias[l], c->stream));
    }
    return DNS_OK;
}

}  // namespace

// ---- internal entry points shared with the closure step (closure_internal.h) -----------
dns_status cnn_forward_device(dns_cnn *c, int B, int n, hipStream_t s)
{
    const int P = B * n * n;
    CHIP(cnn_launch_conv(8, 32, CONV_BIAS_GELU, c->A0, c->Wf[0], c->bias[0], n, P, c->Z1, c->A1, nullptr, 32, s));
    CHIP(cnn_launch_conv(32, 32, CONV_BIAS_GELU, c->A1, c->Wf[1], c->bias[1], n, P, c->Z2, c->A2, nullptr, 32, s));
    CHIP(cnn_launch_conv(32, 32, CONV_BIAS_GELU, c->A2, c->Wf[2], c->bias[2], n, P, c->Z3, c->A3, nullptr, 32, s));
    CHIP(cnn_launch_conv(32, 16, CONV_BIAS, c->A3, c->Wf[3], c->bias[3], n, P, c->Y, nullptr, nullptr, 1, s));
    return DNS_OK;
}

uint16_t *cnn_input_buffer(dns_cnn *c) { return c->A0; }
float *cnn_output_buffer(dns_cnn *c) { return c->Y; }
dns_status cnn_reserve(dns_cnn *c, size_t P) { return ensure_capacity(c, P); }
hipStream_t cnn_stream(dns_cnn *c) { return c->stream; }
int cnn_device(dns_cnn *c) { return c->device; }
const char *cnn_error(dns_cnn *c) { return c->err.c_str(); }

namespace {

// MSE against c->T and the gradient into c->g (forward already done); returns the loss
dns_status backward(dns_cnn *c, int B, int n, double *loss)
{
    const int P = B * n * n;
    const int mse_blocks = std::min(kMseBlocks, (P + 255) / 256);
    CHIP(cnn_launch_mse(c->Y, c->T, P, c->lossp, mse_blocks, c->dz4, c->stream));
    int ppb = (P + kMaxWgradBlocks - 1) / kMaxWgradBlocks;
    ppb = std::max(256, (ppb + 31) / 32 * 32);
    const int blocks = (P + ppb - 1) / ppb;
    const uint16_t *acts[4] = {c->A0, c->A1, c->A2, c->A3};
    uint16_t *dzs[4] = {c->dz1, c->dz2, c->dz3, c->dz4};
    const float *zs[3] = {c->Z1, c->Z2, c->Z3};
    for (int l = 3; l >= 0; --l) {
        const LayerSpec &s = kLayers[l];
        CHIP(cnn_launch_wgrad(s.cinp, s.coutp, dzs[l], acts[l], n, P, ppb, c->part, c->partb, c->stream));
        CHIP(cnn_launch_wgrad_reduce(s.cinp, s.coutp, s.cin, s.cout, c->part, c->partb, blocks, c->g, layer_offset(l),
                                     c->stream));
        if (l > 0)
            CHIP(cnn_launch_conv(s.coutp, s.cinp, CONV_GELU_GRAD, dzs[l], c->Wb[l], nullptr, n, P, nullptr, dzs[l - 1],
                                 zs[l - 1], s.cinp, c->stream));
    }
    if (loss) {
        std::vector<double> lp(mse_blocks);
        CHIP(hipMemcpyAsync(lp.data(), c->lossp, sizeof(double) * mse_blocks, hipMemcpyDeviceToHost, c->stream));
        CHIP(hipStreamSynchronize(c->stream));
        double sum = 0.0;
        for (double x : lp) sum += x;
        *loss = sum / P;
    }
    return DNS_OK;
}

dns_status upload_batch(dns_cnn *c, const float *x, const float *t, int B, int n)
{
    const size_t P = (size_t)B * n * n;
    CHIP(hipMemcpyAsync(c->xstage, x, sizeof(float) * 3 * P, hipMemcpyHostToDevice, c->stream));
    CHIP(cnn_launch_pack_input(c->xstage, B, n, c->A0, c->stream));
    if (t) CHIP(hipMemcpyAsync(c->T, t, sizeof(float) * P, hipMemcpyHostToDevice, c->stream));
    return DNS_OK;
}

dns_status adam_update(dns_cnn *c, float lr)
{
    const double b1 = 0.9, b2 = 0.999;
    c->step += 1;
    const float c1 = (float)(1.0 / (1.0 - std::pow(b1, c->step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(b2, c->step)));
    CHIP(cnn_launch_adam(c->p, c->g, c->m, c->v, kParams, lr, (float)b1, (float)b2, 1e-8f, c1, c2, c->stream));
    return pack_weights(c);
}

}  // namespace

extern "C" int dns_cnn_param_count(void) { return kParams; }

extern "C" const char *dns_cnn_last_error(const dns_cnn *c) { return c ? c->err.c_str() : g_cnn_create_error.c_str(); }

extern "C" void dns_cnn_destroy(dns_cnn *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_activations(c);
    for (void *q : c->fixed) (void)hipFree(q);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" dns_status dns_cnn_create(int device, uint64_t seed, dns_cnn **out)
{
    if (!out) return cfail(nullptr, DNS_E_ARG, "dns_cnn_create: null argument");
    *out = nullptr;
    int ndev = 0;
    hipError_t x = hipGetDeviceCount(&ndev);
    if (x != hipSuccess || ndev <= 0) return cfail(nullptr, DNS_E_HIP, "no HIP device available (%s)", hipGetErrorString(x));
    if (device < 0 || device >= ndev) return cfail(nullptr, DNS_E_ARG, "device %d out of range", device);
    dns_cnn *c = new dns_cnn();
    c->device = device;
    auto bail = [&](dns_status st) {
        g_cnn_create_error = c->err;
        dns_cnn_destroy(c);
        return st;
    };
    if ((x = hipSetDevice(device)) != hipSuccess) return bail(cfail(c, DNS_E_HIP, "hipSetDevice: %s", hipGetErrorString(x)));
    if ((x = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(cfail(c, DNS_E_HIP, "hipStreamCreate: %s", hipGetErrorString(x)));
    auto fixed = [&](auto **p, size_t count, const char *what) {
        dns_status st = calloc_dev(c, p, count, what);
        if (st == DNS_OK) c->fixed.push_back(*p);
        return st;
    };
    dns_status st;
    if ((st = fixed(&c->p, kParams, "parameters")) != DNS_OK) return bail(st);
    if ((st = fixed(&c->m, kParams, "adam m")) != DNS_OK) return bail(st);
    if ((st = fixed(&c->v, kParams, "adam v")) != DNS_OK) return bail(st);
    if ((st = fixed(&c->g, kParams, "gradient")) != DNS_OK) return bail(st);
    if ((st = fixed(&c->lossp, kMseBlocks, "loss partials")) != DNS_OK) return bail(st);
    for (int l = 0; l < 4; ++l) {
        const LayerSpec &s = kLayers[l];
        const size_t nf = (size_t)s.coutp * kp_of(9 * s.cinp), nb = (size_t)s.cinp * kp_of(9 * s.coutp);
        if ((st = fixed(&c->Wf[l], nf, "weights")) != DNS_OK) return bail(st);
        if ((st = fixed(&c->Wb[l], nb, "transposed weights")) != DNS_OK) return bail(st);
        if ((st = fixed(&c->bias[l], (size_t)s.coutp, "bias")) != DNS_OK) return bail(st);
        if ((x = hipMemset(c->Wf[l], 0, nf * 2)) != hipSuccess || (x = hipMemset(c->Wb[l], 0, nb * 2)) != hipSuccess ||
            (x = hipMemset(c->bias[l], 0, s.coutp * 4)) != hipSuccess)
            return bail(cfail(c, DNS_E_HIP, "weight clear: %s", hipGetErrorString(x)));
    }
    // uniform in +-sqrt(3 / fan_in) (variance 1 / fan_in), biases 0
    std::vector<float> p(kParams, 0.0f);
    uint64_t s = seed;
    for (int l = 0; l < 4; ++l) {
        const LayerSpec &ls = kLayers[l];
        const double a = std::sqrt(3.0 / (9.0 * ls.cin));
        const size_t off = layer_offset(l);
        for (int i = 0; i < ls.cout * ls.cin * 9; ++i) {
            const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
            p[off + i] = (float)((2.0 * u - 1.0) * a);
        }
    }
    *out = c;
    st = dns_cnn_set_params(c, p.data());
    if (st != DNS_OK) {
        *out = nullptr;
        return bail(st);
    }
    return DNS_OK;
}

extern "C" dns_status dns_cnn_get_params(dns_cnn *c, float *host)
{
    if (!c || !host) return DNS_E_ARG;
    CHIP(hipSetDevice(c->device));
    CHIP(hipMemcpyAsync(host, c->p, sizeof(float) * kParams, hipMemcpyDeviceToHost, c->stream));
    CHIP(hipStreamSynchronize(c->stream));
    return DNS_OK;
}

extern "C" dns_status dns_cnn_set_params(dns_cnn *c, const float *host)
{
    if (!c || !host) return DNS_E_ARG;
    CHIP(hipSetDevice(c->device));
    CHIP(hipMemcpyAsync(c->p, host, sizeof(float) * kParams, hipMemcpyHostToDevice, c->stream));
    CHIP(hipMemsetAsync(c->m, 0, sizeof(float) * kParams, c->stream));
    CHIP(hipMemsetAsync(c->v, 0, sizeof(float) * kParams, c->stream));
    c->step = 0;
    CSTEP(pack_weights(c));
    CHIP(hipStreamSynchronize(c->stream));
    return DNS_OK;
}

extern "C" dns_status dns_cnn_forward(dns_cnn *c, const float *x, int B, int n, float *y)
{
    if (!c || !x || !y) return DNS_E_ARG;
    CHIP(hipSetDevice(c->device));
    CSTEP(check_shape(c, B, n));
    CSTEP(upload_batch(c, x, nullptr, B, n));
    CSTEP(cnn_forward_device(c, B, n, c->stream));
    CHIP(hipMemcpyAsync(y, c->Y, sizeof(float) * B * n * n, hipMemcpyDeviceToHost, c->stream));
    CHIP(hipStreamSynchronize(c->stream));
    return DNS_OK;
}

extern "C" dns_status dns_cnn_loss_grad(dns_cnn *c, const float *x, const float *t, int B, int n, double *loss,
                                        float *grad)
{
    if (!c || !x || !t) return DNS_E_ARG;
    CHIP(hipSetDevice(c->device));
    CSTEP(check_shape(c, B, n));
    CSTEP(upload_batch(c, x, t, B, n));
    CSTEP(cnn_forward_device(c, B, n, c->stream));
    double l = 0.0;
    CSTEP(backward(c, B, n, &l));
    if (loss) *loss = l;
    if (grad) {
        CHIP(hipMemcpyAsync(grad, c->g, sizeof(float) * kParams, hipMemcpyDeviceToHost, c->stream));
        CHIP(hipStreamSynchronize(c->stream));
    }
    return DNS_OK;
}

extern "C" dns_status dns_cnn_adam_step(dns_cnn *c, const float *x, const float *t, int B, int n, float lr,
                                        double *loss)
{
    if (!c || !x || !t) return DNS_E_ARG;
    CHIP(hipSetDevice(c->device));
    CSTEP(check_shape(c, B, n));
    CSTEP(upload_batch(c, x, t, B, n));
    CSTEP(cnn_forward_device(c, B, n, c->stream));
    CSTEP(backward(c, B, n, loss));
    CSTEP(adam_update(c, lr));
    CHIP(hipStreamSynchronize(c->stream));
    return DNS_OK;
}

extern "C" dns_status dns_cnn_train_on_ensemble(dns_cnn *c, dns_ens *e, int iterations, int batch, float lr,
                                                uint64_t seed, double *loss_history)
{
    if (!c || !e) return DNS_E_ARG;
    if (iterations < 0 || batch < 1) return cfail(c, DNS_E_ARG, "iterations >= 0 and batch >= 1 required");
    CHIP(hipSetDevice(c->device));
    CSTEP(dns_ens_synchronize(e));
    const int ng = dns_ens_group_count(e);
    std::vector<EnsPairView> views(ng);
    double total = 0.0;
    size_t Pmax = 0;
    for (int g = 0; g < ng; ++g) {
        if (dns_ens_group_pairs(e, g, &views[g]) != 0) return cfail(c, DNS_E_ARG, "ensemble group %d unavailable", g);
        if (views[g].stored <= 0) return cfail(c, DNS_E_STATE, "the ensemble holds no training pairs yet");
        total += views[g].count;
        Pmax = std::max(Pmax, (size_t)batch * views[g].n * views[g].n);
        if (((size_t)batch * views[g].n * views[g].n) % 32 != 0)
            return cfail(c, DNS_E_ARG, "batch x n^2 must be a multiple of 32");
    }
    CSTEP(ensure_capacity(c, Pmax));
    for (int it = 0; it < iterations; ++it) {
        // group by a golden-ratio sequence, in proportion to member counts
        const double u = std::fmod(0.6180339887498949 * (it + 1), 1.0) * total;
        int g = 0;
        double acc = views[0].count;
        while (g + 1 < ng && u >= acc) acc += views[++g].count;
        const EnsPairView &v = views[g];
        CHIP(cnn_launch_gather_pairs(v.pairs, v.count, v.stored, v.n, batch, seed, it, c->A0, c->T, c->stream));
        CSTEP(cnn_forward_device(c, batch, v.n, c->stream));
        double l = 0.0;
        CSTEP(backward(c, batch, v.n, loss_history ? &l : nullptr));
        if (loss_history) loss_history[it] = l;
        CSTEP(adam_update(c, lr));
    }
    CHIP(hipStreamSynchronize(c->stream));
    return DNS_OK;
}
