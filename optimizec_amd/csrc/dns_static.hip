// dns_static.hip -- dns_step_noalloc: the time step for callers that cannot allocate
// device memory (code under HIP stream capture, tight serving loops, harnesses that
// forbid hipMalloc).  It runs the fused path's kernels (FP64) on caller-owned om2 /
// fnm1 and takes every work buffer from file-scope device arrays:
//
//   scratch : the transposed state omT, fnT ([kx][iz]), the four z-inverted fields V and
//             the x-forward product G of one problem with N <= kNoallocMaxN.  Every call on
//             a device reuses that device's copy (calls overlapping on one GPU race).
//             Nothing in it must start at zero: each call writes omT, fnT, V and G whole
//             (transpose, column inverse, row pass) before anything reads them.
//   twtab   : the FP64 twiddle tables W_M^t of every supported M (constant after set-up).
//
// Set-up runs once per device, on the first call with that device current: the device
// addresses of both arrays come from hipGetSymbolAddress (the host-side names of the
// arrays are not device addresses and are never passed to a kernel or a copy), and the
// twiddle tables, built on the host in long double exactly as dns_create builds them, are
// copied in with one synchronous hipMemcpy.  Later calls only enqueue kernels on the
// caller's stream.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/dns_hip.h"
#include "dns_kernels.h"

using dnsfft::cd;

namespace {

constexpr int kNoallocMaxN = DNS_NOALLOC_MAX_N;
// supported base grids (the fused kernels built for H = N / 2 <= kNoallocMaxN / 2)
constexpr int kNs[] = {16, 32, 64, 128, 256, 384, 512};
constexpr int kNumN = sizeof(kNs) / sizeof(kNs[0]);

constexpr size_t elems_for(int N)  // complex values of omT, fnT, V (4 fields), G
{
    return (size_t)2 * N * (N / 2) + (size_t)5 * (3 * N / 2) * (N / 2);
}
constexpr size_t twtab_elems()
{
    size_t s = 0;
    for (int N : kNs) s += (size_t)(3 * N / 2);
    return s;
}
constexpr size_t kScratchBytes = elems_for(kNoallocMaxN) * sizeof(cd);
static_assert(kScratchBytes <= (size_t)DNS_NOALLOC_SCRATCH_BYTES, "scratch size in the header is too small");

}  // namespace

static __device__ __align__(256) unsigned char scratch[DNS_NOALLOC_SCRATCH_BYTES];
static __device__ __align__(256) cd twtab[twtab_elems()];

namespace {

constexpr int kMaxDevices = 64;
struct DevPtrs {
    bool ready = false;
    unsigned char *scratch = nullptr;  // device address of `scratch` on this device
    cd *twtab = nullptr;               // device address of `twtab` on this device
};
DevPtrs g_dev[kMaxDevices];
std::mutex g_mu;

dns_status note(dns_status st, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
dns_status note(dns_status st, const char *fmt, ...)
{
    char buf[384];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    dns_note_error(buf);
    return st;
}

// once per device, with that device current
dns_status setup_device(int dev, DevPtrs **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    DevPtrs &d = g_dev[dev];
    if (!d.ready) {
        void *ps = nullptr, *pt = nullptr;
        hipError_t e = hipGetSymbolAddress(&ps, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return note(DNS_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        e = hipGetSymbolAddress(&pt, HIP_SYMBOL(twtab));
        if (e != hipSuccess) return note(DNS_E_HIP, "hipGetSymbolAddress(twtab): %s", hipGetErrorString(e));
        std::vector<cd> tab(twtab_elems());
        size_t off = 0;
        for (int N : kNs) {
            dns_host_twiddles(3 * N / 2, tab.data() + off);
            off += (size_t)(3 * N / 2);
        }
        e = hipMemcpy(pt, tab.data(), sizeof(cd) * tab.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return note(DNS_E_HIP, "twiddle table upload: %s", hipGetErrorString(e));
        d.scratch = static_cast<unsigned char *>(ps);
        d.twtab = static_cast<cd *>(pt);
        d.ready = true;
    }
    *out = &d;
    return DNS_OK;
}

int n_index(int N)
{
    for (int i = 0; i < kNumN; ++i)
        if (kNs[i] == N) return i;
    return -1;
}

}  // namespace

extern "C" dns_status dns_step_noalloc(dns_scalars *sc, double *om2, double *fnm1, int nsteps, void *stream)
{
    if (!sc || !om2 || !fnm1) return note(DNS_E_ARG, "dns_step_noalloc: null argument");
    if (nsteps < 0) return note(DNS_E_ARG, "dns_step_noalloc: nsteps < 0");
    const int N = sc->N;
    if (N > kNoallocMaxN)
        return note(DNS_E_ARG, "dns_step_noalloc: N=%d exceeds the scratch limit N <= %d (%zu bytes)", N,
                    kNoallocMaxN, (size_t)DNS_NOALLOC_SCRATCH_BYTES);
    const int ni = n_index(N);
    if (ni < 0 || !dns_kernels_supported(N / 2))
        return note(DNS_E_UNSUPPORTED, "dns_step_noalloc: N=%d is not one of 16, 32, 64, 128, 256, 384, 512", N);
    if (!(sc->dt > 0.0)) return note(DNS_E_STATE, "dns_step_noalloc: dt must be positive (run NEXTDT first)");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return note(DNS_E_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return note(DNS_E_ARG, "dns_step_noalloc: device %d out of range", dev);
    DevPtrs *d = nullptr;
    dns_status st = setup_device(dev, &d);
    if (st != DNS_OK) return st;
    if (nsteps == 0) return DNS_OK;

    const int H = N / 2, M = 3 * N / 2;
    const size_t nspec = (size_t)N * H, vs = (size_t)M * H;
    cd *omT = reinterpret_cast<cd *>(d->scratch);
    cd *fnT = omT + nspec;
    cd *V = fnT + nspec;
    cd *G = V + 4 * vs;
    size_t toff = 0;
    for (int i = 0; i < ni; ++i) toff += (size_t)(3 * kNs[i] / 2);
    const cd *tw = d->twtab + toff;
    hipStream_t s = static_cast<hipStream_t>(stream);
    cd *om = reinterpret_cast<cd *>(om2);
    cd *fn = reinterpret_cast<cd *>(fnm1);

    // reference layout [iz][ix] -> kernel layout [kx][iz]
    if ((e = dns_launch_transpose(om, omT, N, H, s)) != hipSuccess ||
        (e = dns_launch_transpose(fn, fnT, N, H, s)) != hipSuccess)
        return note(DNS_E_HIP, "dns_step_noalloc: transpose: %s", hipGetErrorString(e));
    static const int f4[4] = {DNS_F_U, DNS_F_W, DNS_F_OX, DNS_F_OZ};
    LaunchReq r{};
    r.what = LAUNCH_COL;
    r.stream = s;
    r.om = omT;
    r.fn = fnT;
    r.G = G;
    r.V = V;
    r.tw = tw;
    r.twT = tw;
    r.col.do_forward = 0;
    r.col.nfields = 4;
    for (int i = 0; i < 4; ++i) r.col.fields[i] = f4[i];
    r.col.vstride = vs;
    r.col.divxz = 1.0 / (double(M) * double(M));
    if ((e = dns_kernels_launch(H, r)) != hipSuccess)
        return note(DNS_E_HIP, "dns_step_noalloc: column inverse: %s", hipGetErrorString(e));
    const double PI = 3.14159265358979323846;
    double t = sc->t, cnm1 = sc->cnm1;
    for (int k = 0; k < nsteps; ++k) {
        // the loop body of dns_step (mem_cuda.cu:1848-1857) on the same kernels
        const double dt = sc->dt;
        LaunchReq w{};
        w.what = LAUNCH_ROW;
        w.stream = s;
        w.V = V;
        w.vstride = vs;
        w.Gw = G;
        w.tw = tw;
        w.twT = tw;
        w.DX = w.DZ = 2.0 * PI / double(N);
        if ((e = dns_kernels_launch(H, w)) != hipSuccess)
            return note(DNS_E_HIP, "dns_step_noalloc: row pass: %s", hipGetErrorString(e));
        LaunchReq c = r;
        c.col.do_forward = 1;
        c.col.VT = 0.5 * sc->visc * dt;
        c.col.c2 = 0.5 * dt * (2.0 + cnm1);
        c.col.c3 = -0.5 * dt * cnm1;
        if ((e = dns_kernels_launch(H, c)) != hipSuccess)
            return note(DNS_E_HIP, "dns_step_noalloc: column pass: %s", hipGetErrorString(e));
        t += dt;
        cnm1 = sc->cn;
    }
    if ((e = dns_launch_transpose(omT, om, H, N, s)) != hipSuccess ||
        (e = dns_launch_transpose(fnT, fn, H, N, s)) != hipSuccess)
        return note(DNS_E_HIP, "dns_step_noalloc: transpose back: %s", hipGetErrorString(e));
    sc->it += nsteps;
    sc->t = t;
    sc->cnm1 = cnm1;
    return DNS_OK;
}
