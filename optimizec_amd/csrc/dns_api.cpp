// dns_api.cpp -- host side of the C ABI (include/dns_hip.h): state, stream, launch
// sequencing, the reference's dt/CFL logic, diagnostics and restart.bin.
//
// The call order and the scalar logic follow the reference driver:
//   init          mem_cuda.cu:1743-1789 (create, scalar reset, PAO + CALCOM, initial
//                 STEP2A, NEXTDT at it=0)
//   loop body     mem_cuda.cu:1848-1857 (it++, dt_old = dt, STEP2B, STEP3, STEP2A,
//                 t += dt_old; STEP3 ends with cnm1 = cn, step3_cuda.cu:172)
//   NEXTDT        dns_cuda.cu:553-580
//   restart.bin   mem_cuda.cu:1019-1394
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cerrno>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dns_closure.h"
#include "../../include/dns_hip.h"
#include "closure_internal.h"
#include "dns_big.h"
#include "dns_kernels.h"

using dnsfft::cd;

namespace {

thread_local std::string g_create_error;

struct TimedLaunch {
    int kernel;
    hipEvent_t a, b;
};

// timed kernels (fused path) and timed phases (multi-pass path, each one or more launches)
enum {
    KT_ROW = 0,
    KT_COL,
    KT_ROW_CFL,
    KT_BIG_ZINV_A,
    KT_BIG_ZINV_B,
    KT_BIG_XINV_A,
    KT_BIG_XINV_B,
    KT_BIG_XFWD_B,
    KT_BIG_ZFWD_A,
    KT_BIG_ZFWD_B,
    // passes 1-3 for the gradient pair only, on the step after a NEXTDT (half-size launches)
    KT_BIG_ZINV_A_HALF,
    KT_BIG_ZINV_B_HALF,
    KT_BIG_XINV_A_HALF,
    KT_COUNT
};
const char *const kKernelNames[KT_COUNT] = {"k_row",           "k_col",           "k_row_cfl",      "big_zinv_a",
                                            "big_zinv_b",      "big_xinv_a",      "big_xinv_b",     "big_xfwd_b",
                                            "big_zfwd_a",      "big_zfwd_b",      "big_zinv_a_half", "big_zinv_b_half",
                                            "big_xinv_a_half"};

}  // namespace

// W_M^t = e^{-2 pi i t / M} for t in [0, M), evaluated in long double and rounded once
void dns_host_twiddles(int M, cd *out)
{
    for (int t = 0; t < M; ++t) {
        const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)M;
        out[t] = cd{(double)cosl(a), -(double)sinl(a)};
    }
}

// error text of a call that has no handle (dns_last_error(NULL))
void dns_note_error(const char *msg) { g_create_error = msg ? msg : ""; }

struct dns_handle {
    int N = 0, M = 0, H = 0, NK = 0;
    int device = 0;
    int fp32 = 0;               // 1: V, G and the transforms in FP32 (state stays FP64)
    int engine = 0;             // with fp32: line-transform stages 0 vector, 1 BF16x3 / 2 FP16 matrix cores,
                                // 3 vector with V, G stored as FP16 pairs
    unsigned *sexp = nullptr;   // engine 3: the storage exponents, vexp (H) then gexp (M)
    bool big = false;           // multi-pass path (dns_big.hip) instead of the fused kernels
    bool div = false;           // fused FP64 kernels in the u_i u_j form: V = (u, w), G = (D^, C^)
    bool big_uiuj = false;      // multi-pass FP64 step in the u_i u_j form (big_step_uiuj)
    void *BA = nullptr;         // multi-pass work buffers (dns_big.h layouts): Y, X, Qh and the
    void *BB = nullptr;         // diagnostics rows in BA; Z, Yq, W, diagnostics plane in BB
    size_t ba_bytes = 0, bb_bytes = 0;
    size_t hs = 0, ms = 0;      // padded row strides of the H- and M-wide rows
    dns_cnn *closure = nullptr;  // network whose forcing each step adds (fused FP64 path)
    cd *twN = nullptr, *caz = nullptr, *cbz = nullptr, *fext = nullptr;
    float *cfields = nullptr, *chp = nullptr;
    double *cpart = nullptr, *cscale = nullptr;
    double band_lo = 0.0, band_hi = 0.0;  // closure forcing band (0, 0: every mode but the mean)
    int limiter = 1;                       // closure: dissipative limiter on (1) / raw forcing (0)
    double nu_dt_max = 0.5;
    dns_scalars sc{};
    hipStream_t stream = nullptr;
    cd *om = nullptr, *fn = nullptr, *tw = nullptr;
    void *V = nullptr, *G = nullptr;  // transform precision (cd or cf)
    void *twT = nullptr;        // twiddles in the transform precision (== tw in FP64)
    void *twH = nullptr;        // fused path: W_H^j = twT[3 j], j < H, contiguous (transform precision)
    cd *Vd = nullptr;           // diagnostics (FP64): 2 z-inverted fields; host staging (lazy)
    double *phys = nullptr;     // diagnostics: one M x M physical plane (lazy)
    double *cflrow = nullptr;   // per-row CFL maxima (M)
    double *partEZ = nullptr;   // per-column E, Z partials (2 H)
    double *red = nullptr;      // min/max partials
    unsigned long long *redu = nullptr;  // pixel partials
    size_t bytes = 0;
    bool have_ic = false;
    bool v_valid = false;
    // multi-pass, FP64: pass 3's plane 0 (u + i w rows, x pass A) of the current om is in
    // buffer A, left there by the NEXTDT CFL; the next step transforms only the gradient pair
    bool x0_valid = false;
    std::string err;
    int timing_every = 0;  // 0: off; k: every k-th launch of each timed kernel
    int64_t seen[KT_COUNT] = {};
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    int64_t launches[KT_COUNT] = {};
    double total_ms[KT_COUNT] = {};
};

namespace {

constexpr int kRedBlocks = 1024;

dns_status fail(dns_handle *h, dns_status st, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h)
        h->err = buf;
    else
        g_create_error = buf;
    return st;
}

#define HIPCHK(h, expr)                                                                            \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail((h), DNS_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

dns_status dev_alloc(dns_handle *h, void **p, size_t bytes, const char *name)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(h, DNS_E_NOMEM, "hipMalloc(%s, %.3f GiB) failed: %s", name, bytes / 1073741824.0,
                    hipGetErrorString(e));
    }
    h->bytes += bytes;
    return DNS_OK;
}

hipEvent_t take_event(dns_handle *h)
{
    if (!h->event_pool.empty()) {
        hipEvent_t e = h->event_pool.back();
        h->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// resolve recorded launches into the per-kernel totals (synchronizes the stream)
void drain_timing(dns_handle *h)
{
    if (h->pending.empty()) return;
    (void)hipStreamSynchronize(h->stream);
    for (TimedLaunch &t : h->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            h->total_ms[t.kernel] += ms;
            h->launches[t.kernel] += 1;
        }
        h->event_pool.push_back(t.a);
        h->event_pool.push_back(t.b);
    }
    h->pending.clear();
}

TimedLaunch timing_begin(dns_handle *h, int timed_kernel)
{
    TimedLaunch tl{timed_kernel, nullptr, nullptr};
    if (h->timing_every > 0 && timed_kernel >= 0 && h->seen[timed_kernel]++ % h->timing_every == 0) {
        if (h->pending.size() > 4096) drain_timing(h);
        tl.a = take_event(h);
        tl.b = take_event(h);
        if (tl.a && tl.b) (void)hipEventRecord(tl.a, h->stream);
    }
    return tl;
}

void timing_end(dns_handle *h, const TimedLaunch &tl)
{
    if (tl.a && tl.b) {
        (void)hipEventRecord(tl.b, h->stream);
        h->pending.push_back(tl);
    }
}

dns_status launch(dns_handle *h, int timed_kernel, LaunchReq &r)
{
    r.stream = h->stream;
    const TimedLaunch tl = timing_begin(h, timed_kernel);
    hipError_t e = dns_kernels_launch(h->H, r);
    if (e != hipSuccess) return fail(h, DNS_E_HIP, "kernel launch (%d) failed: %s", r.what, hipGetErrorString(e));
    timing_end(h, tl);
    return DNS_OK;
}

// ---- multi-pass path ------------------------------------------------------------
dns_status big_launch(dns_handle *h, int timed_kernel, BigReq &r)
{
    r.stream = h->stream;
    r.H = h->H;
    r.hs = h->hs;
    r.ms = h->ms;
    const TimedLaunch tl = timing_begin(h, timed_kernel);
    hipError_t e = dns_big_launch(r);
    if (e != hipSuccess)
        return fail(h, DNS_E_HIP, "multi-pass kernel launch (%d) failed: %s", r.what, hipGetErrorString(e));
    timing_end(h, tl);
    return DNS_OK;
}

// fields (pairs of DNS_F_*) of the current om on the physical rows: Y in BA, Z in BB
dns_status big_zinv(dns_handle *h, bool fp64, const int *fields, int pairs, bool timed)
{
    BigReq r{};
    r.what = BIG_ZINV_A;
    r.fp32 = fp64 ? 0 : h->fp32;
    r.engine = r.fp32 ? h->engine : 0;
    r.tw = r.fp32 ? h->twT : h->tw;
    r.om = h->om;
    r.b = h->BA;
    r.fields.pairs = pairs;
    for (int i = 0; i < 2 * pairs; ++i) r.fields.f[i] = fields[i];
    dns_status st = big_launch(h, timed ? KT_BIG_ZINV_A : -1, r);
    if (st != DNS_OK) return st;
    r.what = BIG_ZINV_B;
    r.a = h->BA;
    r.b = h->BB;
    return big_launch(h, timed ? KT_BIG_ZINV_B : -1, r);
}

// the multi-pass step in the u_i u_j form (FP64): passes 1-3 for u, w only (skipped when the
// NEXTDT on this om left their plane in buffer A), pass 4 forms D = u^2 - w^2 and C = u w, passes
// 5 and 6 carry both, pass 7 forms FN from both
dns_status big_step_uiuj(dns_handle *h, double dt)
{
    static const int fv[2] = {DNS_F_U, DNS_F_W};
    const bool reuse = h->x0_valid;
    h->x0_valid = false;
    dns_status st;
    BigReq r{};
    r.tw = h->tw;
    r.uiuj = 1;
    if (!reuse) {
        // passes 1-3 for one field pair: the half-size launches of the advective step's reuse
        BigReq z{};
        z.what = BIG_ZINV_A;
        z.tw = h->tw;
        z.om = h->om;
        z.b = h->BA;
        z.fields.pairs = 1;
        z.fields.f[0] = fv[0];
        z.fields.f[1] = fv[1];
        if ((st = big_launch(h, KT_BIG_ZINV_A_HALF, z)) != DNS_OK) return st;
        z.what = BIG_ZINV_B;
        z.a = h->BA;
        z.b = h->BB;
        if ((st = big_launch(h, KT_BIG_ZINV_B_HALF, z)) != DNS_OK) return st;
        r.what = BIG_XINV_A;
        r.planes = 1;
        r.a = h->BB;
        r.b = h->BA;
        if ((st = big_launch(h, KT_BIG_XINV_A_HALF, r)) != DNS_OK) return st;
    }
    r.what = BIG_XINV_B;
    r.a = h->BA;
    r.b = h->BB;
    if ((st = big_launch(h, KT_BIG_XINV_B, r)) != DNS_OK) return st;
    r.what = BIG_XFWD_B;
    r.a = h->BB;
    r.b = h->BA;
    if ((st = big_launch(h, KT_BIG_XFWD_B, r)) != DNS_OK) return st;
    const size_t e = sizeof(cd);
    for (int p = 0; p < 2; ++p) {  // pass 6 per plane: Qh plane p (M / 2 rows) -> W field p
        BigReq z = r;
        z.what = BIG_ZFWD_A;
        z.a = static_cast<const char *>(h->BA) + (size_t)p * (h->M / 2) * h->ms * e;
        z.b = static_cast<char *>(h->BB) + (size_t)p * h->M * h->hs * e;
        if ((st = big_launch(h, KT_BIG_ZFWD_A, z)) != DNS_OK) return st;
    }
    r.what = BIG_ZFWD_B;
    r.a = h->BB;
    r.om = h->om;
    r.fn = h->fn;
    r.cn.divxz = 1.0 / (double(h->M) * double(h->M));
    r.cn.VT = 0.5 * h->sc.visc * dt;
    r.cn.c2 = 0.5 * dt * (2.0 + h->sc.cnm1);
    r.cn.c3 = -0.5 * dt * h->sc.cnm1;
    return big_launch(h, KT_BIG_ZFWD_B, r);
}

// one time step of the multi-pass path (seven passes, dns_big.hip)
dns_status big_step(dns_handle *h, double dt)
{
    if (h->big_uiuj) return big_step_uiuj(h, dt);
    static const int f4[4] = {DNS_F_U, DNS_F_W, DNS_F_OX, DNS_F_OZ};
    static const int fg[2] = {DNS_F_OX, DNS_F_OZ};
    const bool reuse = h->x0_valid && !h->fp32;
    h->x0_valid = false;
    dns_status st;
    BigReq r{};
    r.fp32 = h->fp32;
    r.engine = h->engine;
    r.tw = h->fp32 ? h->twT : h->tw;
    if (reuse) {
        // plane 0 of pass 3 (u + i w) is in buffer A from the NEXTDT on this om (same
        // kernels, same arithmetic): passes 1-3 for the gradient pair only, its z-pass-A
        // output in buffer A past plane 0, its rows at their place (fields 2, 3) in buffer B,
        // its plane 1 of pass 3 at its place in buffer A.  Timed under their own names
        // (half-size launches).
        char *ba = static_cast<char *>(h->BA), *bb = static_cast<char *>(h->BB);
        const size_t e = sizeof(cd), plane = (size_t)h->M * h->ms * e, field = (size_t)h->M * h->hs * e;
        BigReq z{};
        z.what = BIG_ZINV_A;
        z.fp32 = 0;
        z.tw = h->tw;
        z.om = h->om;
        z.b = ba + plane;
        z.fields.pairs = 1;
        z.fields.f[0] = fg[0];
        z.fields.f[1] = fg[1];
        if ((st = big_launch(h, KT_BIG_ZINV_A_HALF, z)) != DNS_OK) return st;
        z.what = BIG_ZINV_B;
        z.a = ba + plane;
        z.b = bb + 2 * field;
        if ((st = big_launch(h, KT_BIG_ZINV_B_HALF, z)) != DNS_OK) return st;
        r.what = BIG_XINV_A;
        r.planes = 1;
        r.a = bb + 2 * field;
        r.b = ba + plane;
        if ((st = big_launch(h, KT_BIG_XINV_A_HALF, r)) != DNS_OK) return st;
    } else {
        if ((st = big_zinv(h, false, f4, 2, true)) != DNS_OK) return st;
        r.what = BIG_XINV_A;
        r.planes = 2;
        r.a = h->BB;
        r.b = h->BA;
        if ((st = big_launch(h, KT_BIG_XINV_A, r)) != DNS_OK) return st;
    }
    r.what = BIG_XINV_B;
    r.a = h->BA;
    r.b = h->BB;
    if ((st = big_launch(h, KT_BIG_XINV_B, r)) != DNS_OK) return st;
    r.what = BIG_XFWD_B;
    r.a = h->BB;
    r.b = h->BA;
    if ((st = big_launch(h, KT_BIG_XFWD_B, r)) != DNS_OK) return st;
    r.what = BIG_ZFWD_A;
    r.a = h->BA;
    r.b = h->BB;
    if ((st = big_launch(h, KT_BIG_ZFWD_A, r)) != DNS_OK) return st;
    r.what = BIG_ZFWD_B;
    r.a = h->BB;
    r.om = h->om;
    r.fn = h->fn;
    r.cn.divxz = 1.0 / (double(h->M) * double(h->M));
    r.cn.VT = 0.5 * h->sc.visc * dt;
    r.cn.c2 = 0.5 * dt * (2.0 + h->sc.cnm1);
    r.cn.c3 = -0.5 * dt * h->sc.cnm1;
    return big_launch(h, KT_BIG_ZFWD_B, r);
}

// FP64 physical plane of one field (or |u + i w|) into h->phys, natural order:
// fields on rows (Z in BB), rows R = a (+ i b) in BA, inverse line FFTs (scratch BB)
dns_status big_field(dns_handle *h, int field)
{
    h->x0_valid = false;  // buffers A and B are reused below
    int f2[2];
    bool two = false;
    switch (field) {
        case DNS_FIELD_U: f2[0] = f2[1] = DNS_F_U; break;
        case DNS_FIELD_W: f2[0] = f2[1] = DNS_F_W; break;
        case DNS_FIELD_OMEGA: f2[0] = f2[1] = DNS_F_OMEGA; break;
        case DNS_FIELD_PSI: f2[0] = f2[1] = DNS_F_PSI; break;
        case DNS_FIELD_KINETIC:
            f2[0] = DNS_F_U;
            f2[1] = DNS_F_W;
            two = true;
            break;
        default: return fail(h, DNS_E_ARG, "unknown field %d", field);
    }
    dns_status st = big_zinv(h, true, f2, 1, false);
    if (st != DNS_OK) return st;
    BigReq r{};
    r.what = BIG_ROWS;
    r.tw = h->tw;
    r.a = h->BB;
    r.b = h->BA;
    r.planes = two ? 2 : 1;
    if ((st = big_launch(h, -1, r)) != DNS_OK) return st;
    const size_t line_bytes = sizeof(cd) * (size_t)h->M;
    size_t chunk = h->bb_bytes / line_bytes;
    if (chunk > 65535) chunk = 65535;
    for (size_t l0 = 0; l0 < (size_t)h->M; l0 += chunk) {
        BigReq l{};
        l.what = BIG_LINE_INV;
        l.tw = h->tw;
        l.a = static_cast<char *>(h->BA) + l0 * line_bytes;
        l.b = static_cast<char *>(h->BA) + l0 * line_bytes;
        l.scratch = h->BB;
        l.lines = std::min(chunk, (size_t)h->M - l0);
        if ((st = big_launch(h, -1, l)) != DNS_OK) return st;
    }
    BigReq p{};
    p.what = BIG_ROWS_PHYS;
    p.a = h->BA;
    p.out0 = h->phys;
    p.out_mode = two ? 1 : 0;
    return big_launch(h, -1, p);
}

// CFLM in FP64 through passes 1-3 and the CFL form of pass 4 (any x order)
dns_status big_cflm(dns_handle *h, double *cflm)
{
    const double PI = 3.14159265358979323846;
    static const int f2[2] = {DNS_F_U, DNS_F_W};
    dns_status st = big_zinv(h, true, f2, 1, false);
    if (st != DNS_OK) return st;
    BigReq r{};
    r.tw = h->tw;
    r.what = BIG_XINV_A;
    r.planes = 1;
    r.a = h->BB;
    r.b = h->BA;
    if ((st = big_launch(h, -1, r)) != DNS_OK) return st;
    r.what = BIG_XINV_CFL;
    r.a = h->BA;
    r.cflrow = h->cflrow;
    r.DX = r.DZ = 2.0 * PI / double(h->N);
    if ((st = big_launch(h, -1, r)) != DNS_OK) return st;
    const size_t n = (size_t)dns_big_cfl_parts(h->M);
    std::vector<double> v(n);
    HIPCHK(h, hipMemcpyAsync(v.data(), h->cflrow, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double m = 0.0;
    for (double x : v) m = x > m ? x : m;
    *cflm = m;
    h->x0_valid = !h->fp32;  // plane 0 of pass 3 for this om stays in buffer A
    return DNS_OK;
}

ColArgs inverse_args(dns_handle *h, int nf, const int *fields, size_t vstride)
{
    ColArgs a{};
    a.do_forward = 0;
    a.nfields = nf;
    for (int i = 0; i < nf; ++i) a.fields[i] = fields[i];
    a.vstride = vstride;
    a.divxz = 1.0 / (double(h->M) * double(h->M));
    return a;
}

// the z-inverted fields the step's row kernel reads: u, w (u_i u_j form) or u, w,
// d_x omega, d_z omega (advective form)
int step_fields(const dns_handle *h) { return h->div ? 2 : 4; }

// the FP16 storage exponents of V (per mode) and G (per row), engine 3
void set_storage_exponents(dns_handle *h, LaunchReq &r)
{
    if (h->engine != 3) return;
    r.vexp = h->sexp;
    r.gexp = reinterpret_cast<int *>(h->sexp + h->H);
}

// z-inverse of the step's four spectra (u, w, d_x omega, d_z omega) of the current om
dns_status refresh_v(dns_handle *h)
{
    static const int f4[4] = {DNS_F_U, DNS_F_W, DNS_F_OX, DNS_F_OZ};
    if (h->big) {  // the multi-pass step rebuilds every intermediate from om
        h->v_valid = true;
        return DNS_OK;
    }
    LaunchReq r{};
    r.what = LAUNCH_COL;
    r.fp32 = h->fp32;
    r.engine = h->engine;
    r.om = h->om;
    r.fn = h->fn;
    r.G = h->G;
    r.V = h->V;
    r.tw = h->tw;
    r.twT = h->twT;
    r.twH = h->twH;
    r.col = inverse_args(h, step_fields(h), f4, (size_t)h->M * h->H);
    set_storage_exponents(h, r);
    dns_status st = launch(h, -1, r);
    if (st == DNS_OK) h->v_valid = true;
    return st;
}

dns_status ensure_vd(dns_handle *h)
{
    if (!h->Vd) return dev_alloc(h, (void **)&h->Vd, sizeof(cd) * 2 * (size_t)h->M * h->H, "diag spectra");
    return DNS_OK;
}

dns_status ensure_diag_buffers(dns_handle *h)
{
    dns_status st = ensure_vd(h);
    if (st != DNS_OK) return st;
    if (!h->phys) {
        dns_status st = dev_alloc(h, (void **)&h->phys, sizeof(double) * (size_t)h->M * h->M, "physical plane");
        if (st != DNS_OK) return st;
    }
    return DNS_OK;
}

// physical plane of a dns_field into h->phys
dns_status compute_field(dns_handle *h, int field)
{
    dns_status st = ensure_diag_buffers(h);
    if (st != DNS_OK) return st;
    int fields[2];
    int nf = 1;
    switch (field) {
        case DNS_FIELD_U: fields[0] = DNS_F_U; break;
        case DNS_FIELD_W: fields[0] = DNS_F_W; break;
        case DNS_FIELD_OMEGA: fields[0] = DNS_F_OMEGA; break;
        case DNS_FIELD_PSI: fields[0] = DNS_F_PSI; break;
        case DNS_FIELD_KINETIC:
            fields[0] = DNS_F_U;
            fields[1] = DNS_F_W;
            nf = 2;
            break;
        default: return fail(h, DNS_E_ARG, "unknown field %d", field);
    }
    if (h->big) return big_field(h, field);
    const size_t vs = (size_t)h->M * h->H;
    LaunchReq r{};
    r.what = LAUNCH_COL;
    r.fp32 = 0;  // diagnostics run in FP64 in every precision mode
    r.om = h->om;
    r.fn = h->fn;
    r.G = nullptr;
    r.V = h->Vd;
    r.tw = h->tw;
    r.twT = h->tw;
    r.col = inverse_args(h, nf, fields, vs);
    st = launch(h, -1, r);
    if (st != DNS_OK) return st;
    LaunchReq p{};
    p.what = LAUNCH_ROW_PHYS;
    p.fp32 = 0;
    p.tw = h->tw;
    p.Va = h->Vd;
    p.Vb = (nf == 2) ? h->Vd + vs : nullptr;
    p.out0 = h->phys;
    p.out1 = nullptr;
    p.out_mode = (field == DNS_FIELD_KINETIC) ? 1 : 0;
    return launch(h, -1, p);
}

// CFLM of the current state.  FP64 mode reuses the step's z-inverted u, w (V); in
// FP32 mode u, w are z-inverted again in FP64 into the diagnostics buffer, so the
// CFL/dt logic always sees FP64 velocities.
dns_status compute_cflm(dns_handle *h, double *cflm)
{
    const double PI = 3.14159265358979323846;
    const size_t vs = (size_t)h->M * h->H;
    LaunchReq r{};
    r.what = LAUNCH_ROW_CFL;
    r.fp32 = 0;
    r.vstride = vs;
    r.Gw = nullptr;
    r.tw = h->tw;
    r.twT = h->tw;
    r.cflrow = h->cflrow;
    r.DX = 2.0 * PI / double(h->N);
    r.DZ = 2.0 * PI / double(h->N);
    if (h->big) return big_cflm(h, cflm);
    if (!h->fp32) {
        if (!h->v_valid) {
            dns_status st = refresh_v(h);
            if (st != DNS_OK) return st;
        }
        r.V = h->V;
    } else {
        dns_status st = ensure_vd(h);
        if (st != DNS_OK) return st;
        static const int f2[2] = {DNS_F_U, DNS_F_W};
        LaunchReq c{};
        c.what = LAUNCH_COL;
        c.fp32 = 0;
        c.om = h->om;
        c.fn = h->fn;
        c.V = h->Vd;
        c.tw = h->tw;
        c.twT = h->tw;
        c.col = inverse_args(h, 2, f2, vs);
        st = launch(h, -1, c);
        if (st != DNS_OK) return st;
        r.V = h->Vd;
    }
    dns_status st = launch(h, KT_ROW_CFL, r);
    if (st != DNS_OK) return st;
    std::vector<double> rows(h->M);
    HIPCHK(h, hipMemcpyAsync(rows.data(), h->cflrow, sizeof(double) * h->M, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double m = 0.0;
    for (double v : rows) m = v > m ? v : m;
    *cflm = m;
    return DNS_OK;
}

// elements of om (and of fn) on the device
size_t state_elems(const dns_handle *h)
{
    return h->big ? (size_t)h->N * 8 * dns_state_tiles(h->H) : (size_t)h->N * h->H;
}

// host [iz][ix] <-> device [ix][iz] through the diagnostics buffer as staging (fused path);
// host rows <-> the state's column tiles through buffer B (multi-pass path; B holds nothing
// that outlives a step)
dns_status upload_spec(dns_handle *h, const double *host, cd *dst)
{
    const size_t n = (size_t)h->N * h->H;
    h->x0_valid = false;  // the state changes
    if (h->big) {
        HIPCHK(h, hipMemcpyAsync(h->BB, host, sizeof(cd) * n, hipMemcpyHostToDevice, h->stream));
        BigReq r{};
        r.what = BIG_STATE_TILE;
        r.a = h->BB;
        r.b = dst;
        dns_status st = big_launch(h, -1, r);
        if (st != DNS_OK) return st;
        HIPCHK(h, hipStreamSynchronize(h->stream));
        return DNS_OK;
    }
    dns_status st = ensure_vd(h);
    if (st != DNS_OK) return st;
    HIPCHK(h, hipMemcpyAsync(h->Vd, host, sizeof(cd) * n, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, dns_launch_transpose(h->Vd, dst, h->N, h->H, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DNS_OK;
}

dns_status download_spec(dns_handle *h, const cd *src, double *host)
{
    const size_t n = (size_t)h->N * h->H;
    if (h->big) {
        BigReq r{};
        r.what = BIG_STATE_UNTILE;
        r.a = src;
        r.b = h->BB;
        dns_status st = big_launch(h, -1, r);
        if (st != DNS_OK) return st;
        HIPCHK(h, hipMemcpyAsync(host, h->BB, sizeof(cd) * n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        return DNS_OK;
    }
    dns_status st = ensure_vd(h);
    if (st != DNS_OK) return st;
    HIPCHK(h, dns_launch_transpose(src, h->Vd, h->H, h->N, h->stream));
    HIPCHK(h, hipMemcpyAsync(host, h->Vd, sizeof(cd) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DNS_OK;
}

// subgrid forcing of the attached network for the current om into h->fext
dns_status closure_forcing(dns_handle *h)
{
    const int N = h->N;
    HIPCHK(h, closure_launch_to_grid(h->om, N, h->twN, h->caz, h->cfields, h->stream));
    HIPCHK(h, closure_launch_pack(h->cfields, N, h->chp, h->cpart, h->cscale, cnn_input_buffer(h->closure),
                                  h->stream));
    if (cnn_forward_device(h->closure, 1, N, h->stream) != DNS_OK)
        return fail(h, DNS_E_HIP, "closure network: %s", cnn_error(h->closure));
    ClosureShape sh;
    sh.k_lo = h->band_lo;
    sh.k_hi = h->band_hi;
    if (h->limiter) {
        sh.om = h->om;
        sh.dt = h->sc.dt;
        sh.nu_dt_max = h->nu_dt_max;
    }
    HIPCHK(h, closure_launch_from_grid(cnn_output_buffer(h->closure), h->cscale, N, h->twN, h->cbz, h->fext, sh,
                                       h->stream));
    return DNS_OK;
}

void reset_scalars(dns_handle *h)
{
    // mem_cuda.cu:1755-1765
    h->sc.it = 0;
    h->sc.ifn = 1;
    h->sc.t = 0.0;
    h->sc.dt = 0.0;
    h->sc.cn = 1.0;
    h->sc.cnm1 = 0.0;
    h->sc.cflm = 0.0;
}

}  // namespace

// =====================================================================================
// life cycle
// =====================================================================================
extern "C" int dns_supported_path(int N, int path)
{
    if (N < 16 || N % 4 != 0) return 0;
    const bool fused = dns_kernels_supported(N / 2) != 0, big = dns_big_supported_m(3 * N / 2) != 0;
    switch (path) {
        case DNS_PATH_AUTO: return (fused || big) ? 1 : 0;
        case DNS_PATH_FUSED: return fused ? 1 : 0;
        case DNS_PATH_MULTIPASS: return big ? 1 : 0;
        case DNS_PATH_FUSED_UIUJ: return (fused && dns_kernels_uiuj_supported(N / 2)) ? 1 : 0;
        case DNS_PATH_MULTIPASS_UIUJ: return big ? 1 : 0;
        default: return 0;
    }
}

extern "C" int dns_supported_n(int N) { return dns_supported_path(N, DNS_PATH_AUTO); }

extern "C" const char *dns_last_error(const dns_handle *h)
{
    return h ? h->err.c_str() : g_create_error.c_str();
}

extern "C" dns_status dns_create(const dns_config *cfg, dns_handle **out)
{
    if (!cfg || !out) return fail(nullptr, DNS_E_ARG, "dns_create: null argument");
    *out = nullptr;
    if (cfg->N < 4 || cfg->N % 4 != 0) return fail(nullptr, DNS_E_ARG, "N=%d must be a positive multiple of 4", cfg->N);
    if (cfg->path < DNS_PATH_AUTO || cfg->path > DNS_PATH_MULTIPASS_UIUJ)
        return fail(nullptr, DNS_E_ARG, "unknown path %d", cfg->path);
    if (!dns_supported_path(cfg->N, cfg->path))
        return fail(nullptr, DNS_E_UNSUPPORTED, "N=%d: no kernels built for H=N/2=%d, M=3N/2=%d on path %d", cfg->N,
                    cfg->N / 2, 3 * cfg->N / 2, cfg->path);
    if (cfg->precision < DNS_PREC_FP64 || cfg->precision > DNS_PREC_FP16_STORE)
        return fail(nullptr, DNS_E_UNSUPPORTED, "precision %d is not built", cfg->precision);
    if (cfg->precision != DNS_PREC_FP64 && cfg->N < 128)
        return fail(nullptr, DNS_E_UNSUPPORTED, "FP32 transforms are built for N >= 128");
    if (cfg->precision == DNS_PREC_BF16X3 || cfg->precision == DNS_PREC_FP16) {
        const bool fused = cfg->path != DNS_PATH_MULTIPASS && dns_kernels_supported(cfg->N / 2);
        const bool ok = fused ? dns_kernels_mfma_supported(cfg->N / 2) != 0 : dns_big_mfma_supported_m(3 * cfg->N / 2) != 0;
        if (!ok)
            return fail(nullptr, DNS_E_UNSUPPORTED, "matrix-core line transforms are not built for N=%d on path %d",
                        cfg->N, cfg->path);
    }
    if (cfg->path == DNS_PATH_MULTIPASS_UIUJ && cfg->precision != DNS_PREC_FP64)
        return fail(nullptr, DNS_E_UNSUPPORTED, "the u_i u_j form of the multi-pass step is built for FP64, not precision %d",
                    cfg->precision);
    if (cfg->path == DNS_PATH_FUSED_UIUJ && (cfg->precision != DNS_PREC_FP64 || !dns_kernels_uiuj_supported(cfg->N / 2)))
        return fail(nullptr, DNS_E_UNSUPPORTED, "the u_i u_j form is built for FP64 at N=4096, not N=%d precision %d",
                    cfg->N, cfg->precision);
    if (cfg->precision == DNS_PREC_FP16_STORE &&
        (cfg->path == DNS_PATH_MULTIPASS || !dns_kernels_supported(cfg->N / 2) || !dns_kernels_h16_supported(cfg->N / 2)))
        return fail(nullptr, DNS_E_UNSUPPORTED, "FP16 storage of the intermediates is built for N=4096 on the fused path, not N=%d path %d",
                    cfg->N, cfg->path);
    if (!(cfg->Re > 0.0)) return fail(nullptr, DNS_E_ARG, "Re must be positive");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0)
        return fail(nullptr, DNS_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (cfg->device < 0 || cfg->device >= ndev) return fail(nullptr, DNS_E_ARG, "device %d out of range", cfg->device);

    dns_handle *h = new dns_handle();
    h->N = cfg->N;
    h->H = cfg->N / 2;
    h->M = 3 * cfg->N / 2;
    h->NK = 3 * cfg->N / 4 + 1;
    h->device = cfg->device;
    h->fp32 = cfg->precision != DNS_PREC_FP64 ? 1 : 0;
    h->engine = cfg->precision == DNS_PREC_BF16X3       ? 1
                : cfg->precision == DNS_PREC_FP16       ? 2
                : cfg->precision == DNS_PREC_FP16_STORE ? 3
                                                        : 0;
    h->big = cfg->path == DNS_PATH_MULTIPASS || cfg->path == DNS_PATH_MULTIPASS_UIUJ ||
             (cfg->path == DNS_PATH_AUTO && !dns_kernels_supported(h->H));
    h->big_uiuj = cfg->path == DNS_PATH_MULTIPASS_UIUJ;
    h->sc.N = h->N;
    h->sc.M = h->M;
    h->sc.H = h->H;
    h->sc.NK = h->NK;
    h->sc.Re = cfg->Re;
    h->sc.K0 = cfg->K0;
    h->sc.visc = 0.0;
    h->sc.cflnum = cfg->cflnum;
    reset_scalars(h);

    auto bail = [&](dns_status st) {
        g_create_error = h->err;
        dns_destroy(h);
        return st;
    };
    if ((e = hipSetDevice(h->device)) != hipSuccess)
        return bail(fail(h, DNS_E_HIP, "hipSetDevice: %s", hipGetErrorString(e)));
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(fail(h, DNS_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e)));

    // the multi-pass state is in column tiles of 8 modes (the last one padded)
    const size_t nspec = state_elems(h);
    const size_t nmid = (size_t)h->M * h->H;
    dns_status st;
    if ((st = dev_alloc(h, (void **)&h->om, sizeof(cd) * nspec, "om2")) != DNS_OK) return bail(st);
    if ((st = dev_alloc(h, (void **)&h->fn, sizeof(cd) * nspec, "fnm1")) != DNS_OK) return bail(st);
    // bytes of one stored V / G value (FP16 pairs with engine 3)
    const size_t elem = h->engine == 3 ? 4 : h->fp32 ? sizeof(dnsfft::cf) : sizeof(cd);
    if (h->big) {
        // BA holds Y (4 fields), X (2 planes), Qh; BB holds Z (4 fields), Yq, W.  The FP64
        // diagnostics use the same bytes: Y, X-plane 0 and rows R in BA; Z, line-FFT
        // scratch and the physical plane in BB.
        const size_t M = h->M;
        h->hs = ((size_t)h->H + 15) / 16 * 16;
        h->ms = (M + 15) / 16 * 16;
        const size_t ba = std::max({elem * 2 * M * h->ms, elem * 4 * M * h->hs, sizeof(cd) * 2 * M * h->hs,
                                    sizeof(cd) * M * h->ms, sizeof(cd) * M * M});
        const size_t bb = std::max({elem * 4 * M * h->hs, elem * (M / 2) * h->ms, sizeof(cd) * 2 * M * h->hs,
                                    sizeof(double) * M * M, sizeof(cd) * M * 8});
        h->ba_bytes = ba;
        h->bb_bytes = bb;
        if ((st = dev_alloc(h, &h->BA, ba, "multi-pass buffer A")) != DNS_OK) return bail(st);
        if ((st = dev_alloc(h, &h->BB, bb, "multi-pass buffer B")) != DNS_OK) return bail(st);
        h->phys = static_cast<double *>(h->BB);
    } else {
        // V: the step's z-inverted fields; G: the x-forward products (D^, C^ in the u_i u_j
        // form, the one advective product otherwise)
        h->div = cfg->path == DNS_PATH_FUSED_UIUJ;
        if ((st = dev_alloc(h, &h->V, elem * step_fields(h) * nmid, "z-inverse fields")) != DNS_OK) return bail(st);
        if ((st = dev_alloc(h, &h->G, elem * (h->div ? 2 : 1) * nmid, "x-forward products")) != DNS_OK)
            return bail(st);
        if (h->engine == 3 &&
            (st = dev_alloc(h, (void **)&h->sexp, sizeof(unsigned) * (size_t)(h->H + h->M), "storage exponents")) != DNS_OK)
            return bail(st);
    }
    if ((st = dev_alloc(h, (void **)&h->tw, sizeof(cd) * h->M, "twiddles")) != DNS_OK) return bail(st);
    const size_t ncfl = h->big ? (size_t)dns_big_cfl_parts(h->M) : (size_t)h->M;
    if ((st = dev_alloc(h, (void **)&h->cflrow, sizeof(double) * ncfl, "cfl maxima")) != DNS_OK) return bail(st);
    if ((st = dev_alloc(h, (void **)&h->partEZ, sizeof(double) * 2 * h->N, "energy partials")) != DNS_OK)
        return bail(st);
    if ((st = dev_alloc(h, (void **)&h->red, sizeof(double) * 2 * kRedBlocks, "minmax partials")) != DNS_OK)
        return bail(st);
    if ((st = dev_alloc(h, (void **)&h->redu, sizeof(unsigned long long) * 2 * kRedBlocks, "pixel partials")) !=
        DNS_OK)
        return bail(st);

    std::vector<cd> tw(h->M);
    dns_host_twiddles(h->M, tw.data());
    if ((e = hipMemcpy(h->tw, tw.data(), sizeof(cd) * h->M, hipMemcpyHostToDevice)) != hipSuccess)
        return bail(fail(h, DNS_E_HIP, "twiddle upload: %s", hipGetErrorString(e)));
    h->twT = h->tw;
    if (h->fp32) {
        std::vector<dnsfft::cf> twf(h->M);
        for (int t = 0; t < h->M; ++t) twf[t] = dnsfft::cf{float(tw[t].x), float(tw[t].y)};
        if ((st = dev_alloc(h, &h->twT, sizeof(dnsfft::cf) * h->M, "fp32 twiddles")) != DNS_OK) return bail(st);
        if ((e = hipMemcpy(h->twT, twf.data(), sizeof(dnsfft::cf) * h->M, hipMemcpyHostToDevice)) != hipSuccess)
            return bail(fail(h, DNS_E_HIP, "twiddle upload: %s", hipGetErrorString(e)));
    }
    if (!h->big) {
        const size_t ebytes = h->fp32 ? sizeof(dnsfft::cf) : sizeof(cd);
        std::vector<unsigned char> th(ebytes * (size_t)h->H);
        for (int j = 0; j < h->H; ++j) {
            if (h->fp32) {
                const dnsfft::cf v{float(tw[3 * j].x), float(tw[3 * j].y)};
                std::memcpy(th.data() + ebytes * j, &v, ebytes);
            } else {
                std::memcpy(th.data() + ebytes * j, &tw[3 * j], ebytes);
            }
        }
        if ((st = dev_alloc(h, &h->twH, th.size(), "W_H twiddles")) != DNS_OK) return bail(st);
        if ((e = hipMemcpy(h->twH, th.data(), th.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return bail(fail(h, DNS_E_HIP, "twiddle upload: %s", hipGetErrorString(e)));
    }
    if ((e = hipMemset(h->om, 0, sizeof(cd) * nspec)) != hipSuccess ||
        (e = hipMemset(h->fn, 0, sizeof(cd) * nspec)) != hipSuccess)
        return bail(fail(h, DNS_E_HIP, "state clear: %s", hipGetErrorString(e)));
    *out = h;
    return DNS_OK;
}

extern "C" void dns_destroy(dns_handle *h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (TimedLaunch &t : h->pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (hipEvent_t ev : h->event_pool) (void)hipEventDestroy(ev);
    if (h->twT == h->tw) h->twT = nullptr;
    if (h->big) h->phys = nullptr;  // a view of buffer B
    void *ptrs[] = {h->om,  h->fn,     h->V,   h->G,    h->tw,  h->twT, h->twH, h->Vd,   h->phys,  h->cflrow,
                    h->partEZ, h->red, h->redu, h->BA,   h->BB,  nullptr,  h->twN, h->caz, h->cbz,
                    h->fext, h->cfields, h->chp, h->cpart, h->cscale, h->sexp};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

// =====================================================================================
// initial condition and stepping
// =====================================================================================
extern "C" dns_status dns_init_ic(dns_handle *h, int kind, uint64_t seed, char *log, int logcap)
{
    if (!h) return DNS_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    std::vector<double> om2((size_t)2 * h->N * h->H);
    double visc = 1.0 / h->sc.Re;
    if (kind == DNS_IC_PAO_KM3) {
        if (dns_ic_pao_host(h->N, h->sc.Re, h->sc.K0, om2.data(), log, logcap, &visc) < 0)
            return fail(h, DNS_E_ARG, "PAO initial condition rejected its arguments");
    } else if (kind == DNS_IC_K4EXP) {
        if (dns_ic_k4exp_host(h->N, h->sc.K0, seed, om2.data()) != DNS_OK)
            return fail(h, DNS_E_ARG, "k4exp initial condition rejected its arguments (K0 > 0?)");
        if (log && logcap > 0) log[0] = '\0';
    } else {
        return fail(h, DNS_E_ARG, "unknown initial condition kind %d", kind);
    }
    reset_scalars(h);
    h->sc.visc = visc;
    dns_status st = upload_spec(h, om2.data(), h->om);
    if (st != DNS_OK) return st;
    HIPCHK(h, hipMemsetAsync(h->fn, 0, sizeof(cd) * state_elems(h), h->stream));
    h->have_ic = true;
    return refresh_v(h);
}

extern "C" dns_status dns_start(dns_handle *h)
{
    if (!h) return DNS_E_ARG;
    if (!h->have_ic) return fail(h, DNS_E_STATE, "dns_start before an initial condition");
    if (!(h->sc.it == 0 && h->sc.dt == 0.0)) return fail(h, DNS_E_STATE, "dns_start needs it == 0 and dt == 0");
    return dns_next_dt(h);
}

extern "C" dns_status dns_step(dns_handle *h, int nsteps)
{
    if (!h) return DNS_E_ARG;
    if (!h->have_ic) return fail(h, DNS_E_STATE, "dns_step before an initial condition");
    if (nsteps < 0) return fail(h, DNS_E_ARG, "nsteps < 0");
    HIPCHK(h, hipSetDevice(h->device));
    if (!h->v_valid) {
        dns_status st = refresh_v(h);
        if (st != DNS_OK) return st;
    }
    static const int f4[4] = {DNS_F_U, DNS_F_W, DNS_F_OX, DNS_F_OZ};
    const size_t vs = (size_t)h->M * h->H;
    const double PI = 3.14159265358979323846;
    for (int s = 0; s < nsteps; ++s) {
        h->sc.it += 1;
        const double dt = h->sc.dt;
        if (h->big) {
            dns_status st = big_step(h, dt);
            if (st != DNS_OK) return st;
            h->sc.t += dt;
            h->sc.cnm1 = h->sc.cn;
            continue;
        }
        if (h->closure) {
            dns_status st = closure_forcing(h);
            if (st != DNS_OK) return st;
        }
        LaunchReq r{};
        r.what = LAUNCH_ROW;
        r.fp32 = h->fp32;
        r.engine = h->engine;
        r.V = h->V;
        r.vstride = vs;
        r.Gw = h->G;
        r.tw = h->tw;
        r.twT = h->twT;
        r.twH = h->twH;
        r.cflrow = nullptr;
        r.DX = r.DZ = 2.0 * PI / double(h->N);
        r.form = h->div ? 1 : 0;
        set_storage_exponents(h, r);
        dns_status st = launch(h, KT_ROW, r);
        if (st != DNS_OK) return st;
        LaunchReq c{};
        c.what = LAUNCH_COL;
        c.fp32 = h->fp32;
        c.engine = h->engine;
        c.om = h->om;
        c.fn = h->fn;
        c.G = h->G;
        c.V = h->V;
        c.tw = h->tw;
        c.twT = h->twT;
        c.twH = h->twH;
        c.col.do_forward = 1;
        c.col.nfields = step_fields(h);
        for (int i = 0; i < 4; ++i) c.col.fields[i] = f4[i];
        c.col.vstride = vs;
        c.col.divxz = 1.0 / (double(h->M) * double(h->M));
        c.col.VT = 0.5 * h->sc.visc * dt;
        c.col.c2 = 0.5 * dt * (2.0 + h->sc.cnm1);
        c.col.c3 = -0.5 * dt * h->sc.cnm1;
        c.col.fext = h->closure ? h->fext : nullptr;
        c.form = h->div ? 1 : 0;
        set_storage_exponents(h, c);
        st = launch(h, KT_COL, c);
        if (st != DNS_OK) return st;
        h->sc.t += dt;
        h->sc.cnm1 = h->sc.cn;
    }
    return DNS_OK;
}

extern "C" dns_status dns_next_dt(dns_handle *h)
{
    if (!h) return DNS_E_ARG;
    if (!h->have_ic) return fail(h, DNS_E_STATE, "dns_next_dt before an initial condition");
    HIPCHK(h, hipSetDevice(h->device));
    const double PI = 3.14159265358979323846;
    if (h->sc.it == 0 && h->sc.dt == 0.0) {
        double cflm = 0.0;
        dns_status st = compute_cflm(h, &cflm);
        if (st != DNS_OK) return st;
        h->sc.dt = h->sc.cflnum / cflm / PI;
        h->sc.cflm = cflm;
        return DNS_OK;
    }
    if (h->sc.ifn > 0 && (h->sc.it % h->sc.ifn) == 0) {
        double cflm = 0.0;
        dns_status st = compute_cflm(h, &cflm);
        if (st != DNS_OK) return st;
        const double cfl = cflm * h->sc.dt * PI;
        h->sc.cn = 0.8 + 0.2 * (h->sc.cflnum / cfl);
        h->sc.dt = h->sc.dt * h->sc.cn;
        h->sc.cflm = cflm;
    }
    return DNS_OK;
}

// =====================================================================================
// diagnostics
// =====================================================================================
extern "C" dns_status dns_energy_enstrophy(dns_handle *h, double *E, double *Z)
{
    if (!h || !E || !Z) return DNS_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    // per-column partials (fused layout [kx][iz]) or per-tile partials (multi-pass column tiles)
    const int parts = h->big ? dns_state_tiles(h->H) : h->H;
    if (h->big) {
        BigReq r{};
        r.what = BIG_ENERGY;
        r.om = h->om;
        r.out0 = h->partEZ;
        r.out1 = h->partEZ + parts;
        dns_status st = big_launch(h, -1, r);
        if (st != DNS_OK) return st;
    } else {
        HIPCHK(h, dns_launch_energy(h->om, h->N, h->H, h->partEZ, h->partEZ + parts, h->stream));
    }
    std::vector<double> p(2 * (size_t)parts);
    HIPCHK(h, hipMemcpyAsync(p.data(), h->partEZ, sizeof(double) * p.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double e = 0.0, z = 0.0;
    for (int i = 0; i < parts; ++i) {
        e += p[i];
        z += p[parts + i];
    }
    *E = e;
    *Z = z;
    return DNS_OK;
}

// computePalOverEnsKmax2 (mem_cuda.cu:872-935) on the device: per-line partial sums, then
// the ratio on the host; kmax2 = (H - 1)^2 + H^2, the largest K^2 of the stored modes
extern "C" dns_status dns_pal_over_ens_kmax2(dns_handle *h, double *out)
{
    if (!h || !out) return DNS_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    const int parts = h->big ? dns_state_tiles(h->H) : h->H;
    HIPCHK(h, dns_launch_palin(h->om, h->N, h->big ? 1 : 0, h->partEZ, h->partEZ + parts, h->stream));
    std::vector<double> p(2 * (size_t)parts);
    HIPCHK(h, hipMemcpyAsync(p.data(), h->partEZ, sizeof(double) * p.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double total = 0.0, pal = 0.0;
    for (int i = 0; i < parts; ++i) {
        total += p[i];
        pal += p[parts + i];
    }
    const double kmax2 = double(h->H - 1) * double(h->H - 1) + double(h->H) * double(h->H);
    *out = (total <= 0.0) ? 0.0 : pal / (total * kmax2);
    return DNS_OK;
}

extern "C" dns_status dns_display_sigma(dns_handle *h, double *sigma)
{
    if (!h || !sigma) return DNS_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    dns_status st = compute_field(h, DNS_FIELD_OMEGA);
    if (st != DNS_OK) return st;
    const size_t n = (size_t)h->M * h->M;
    const int blocks = (int)std::min<size_t>(kRedBlocks, (n + 255) / 256);
    HIPCHK(h, dns_launch_minmax(h->phys, n, h->red, blocks, h->stream));
    std::vector<double> mm(2 * (size_t)blocks);
    HIPCHK(h, hipMemcpyAsync(mm.data(), h->red, sizeof(double) * mm.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double mn = mm[0], mx = mm[1];
    for (int b = 1; b < blocks; ++b) {
        mn = std::min(mn, mm[2 * b]);
        mx = std::max(mx, mm[2 * b + 1]);
    }
    const double rng = mx - mn;
    const int is_const = std::fabs(rng) <= 1.0e-12 ? 1 : 0;
    HIPCHK(h, dns_launch_pixels(h->phys, n, mn, rng, is_const, h->redu, blocks, h->stream));
    std::vector<unsigned long long> ps(2 * (size_t)blocks);
    HIPCHK(h, hipMemcpyAsync(ps.data(), h->redu, sizeof(unsigned long long) * ps.size(), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    unsigned long long s1 = 0, s2 = 0;
    for (int b = 0; b < blocks; ++b) {
        s1 += ps[2 * b];
        s2 += ps[2 * b + 1];
    }
    const double mean = double(s1) / double(n);
    const double var = std::max(0.0, double(s2) / double(n) - mean * mean);
    *sigma = std::sqrt(var);
    return DNS_OK;
}

extern "C" dns_status dns_fetch_field(dns_handle *h, int field, double *host)
{
    if (!h || !host) return DNS_E_ARG;
    if (!h->have_ic) return fail(h, DNS_E_STATE, "dns_fetch_field before an initial condition");
    HIPCHK(h, hipSetDevice(h->device));
    dns_status st = compute_field(h, field);
    if (st != DNS_OK) return st;
    HIPCHK(h, hipMemcpyAsync(host, h->phys, sizeof(double) * (size_t)h->M * h->M, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DNS_OK;
}

// =====================================================================================
// state access
// =====================================================================================
extern "C" dns_status dns_get_scalars(const dns_handle *h, dns_scalars *out)
{
    if (!h || !out) return DNS_E_ARG;
    *out = h->sc;
    return DNS_OK;
}

extern "C" dns_status dns_set_scalars(dns_handle *h, const dns_scalars *in)
{
    if (!h || !in) return DNS_E_ARG;
    h->sc.it = in->it;
    h->sc.ifn = in->ifn;
    h->sc.Re = in->Re;
    h->sc.K0 = in->K0;
    h->sc.visc = in->visc;
    h->sc.t = in->t;
    h->sc.dt = in->dt;
    h->sc.cn = in->cn;
    h->sc.cnm1 = in->cnm1;
    h->sc.cflnum = in->cflnum;
    h->sc.cflm = in->cflm;
    return DNS_OK;
}

extern "C" dns_status dns_get_om2(dns_handle *h, double *host)
{
    if (!h || !host) return DNS_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    return download_spec(h, h->om, host);
}

extern "C" dns_status dns_set_om2(dns_handle *h, const double *host)
{
    if (!h || !host) return DNS_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    dns_status st = upload_spec(h, host, h->om);
    if (st != DNS_OK) return st;
    h->have_ic = true;
    h->v_valid = false;
    return refresh_v(h);
}

extern "C" dns_status dns_get_fnm1(dns_handle *h, double *host)
{
    if (!h || !host) return DNS_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    return download_spec(h, h->fn, host);
}

extern "C" dns_status dns_set_fnm1(dns_handle *h, const double *host)
{
    if (!h || !host) return DNS_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    return upload_spec(h, host, h->fn);
}

// =====================================================================================
// restart.bin
// =====================================================================================
namespace {

struct HeadInts {
    char magic[16];
    uint32_t version, header_bytes, real_bytes, cplx_bytes;
    int32_t Nbase, N, NX, NZ, NK, NX_full, NZ_full, NK_full, it, ifn;
};
static_assert(sizeof(HeadInts) == 72, "restart header integer block is 72 bytes");

constexpr size_t kScalarCount = 9;  // Re K0 visc t dt cn cnm1 cflnum cflm

size_t align8(size_t x) { return (x + 7) & ~size_t(7); }
// header sizes of RestartHeaderV1 / RestartHeader for a given real size
size_t v1_bytes(size_t real_bytes) { return align8(72 + kScalarCount * real_bytes) + 24; }
size_t v2_bytes(size_t real_bytes) { return v1_bytes(real_bytes) + 8; }

bool write_all(FILE *fp, const void *p, size_t n) { return n == 0 || std::fwrite(p, 1, n, fp) == n; }
bool read_all(FILE *fp, void *p, size_t n) { return n == 0 || std::fread(p, 1, n, fp) == n; }

}  // namespace

extern "C" dns_status dns_save_restart(dns_handle *h, const char *path)
{
    if (!h || !path) return DNS_E_ARG;
    if (!h->have_ic) return fail(h, DNS_E_STATE, "dns_save_restart before an initial condition");
    HIPCHK(h, hipSetDevice(h->device));
    const size_t hb = v2_bytes(8);
    std::vector<unsigned char> head(hb, 0);
    HeadInts hi{};
    std::snprintf(hi.magic, sizeof hi.magic, "CTRB_RESTART");
    hi.version = 2;
    hi.header_bytes = (uint32_t)hb;
    hi.real_bytes = 8;
    hi.cplx_bytes = 16;
    hi.Nbase = hi.N = hi.NX = hi.NZ = h->N;
    hi.NK = hi.NK_full = h->NK;
    hi.NX_full = hi.NZ_full = h->M;
    hi.it = h->sc.it;
    hi.ifn = h->sc.ifn;
    std::memcpy(head.data(), &hi, sizeof hi);
    const double sc[kScalarCount] = {h->sc.Re, h->sc.K0, h->sc.visc, h->sc.t, h->sc.dt,
                                     h->sc.cn, h->sc.cnm1, h->sc.cflnum, h->sc.cflm};
    std::memcpy(head.data() + 72, sc, sizeof sc);
    const uint64_t om_count = (uint64_t)h->H * h->N;
    const uint64_t ur_count = (uint64_t)h->M * h->M * 3;
    const uint64_t counts[4] = {0, om_count, om_count, ur_count};  // uc, om2, fnm1, ur
    std::memcpy(head.data() + align8(72 + sizeof sc), counts, sizeof counts);

    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(h, DNS_E_IO, "cannot open %s for writing: %s", path, std::strerror(errno));
    bool ok = write_all(fp, head.data(), hb);
    std::vector<double> plane((size_t)h->M * h->M);
    const int comps[3] = {DNS_FIELD_U, DNS_FIELD_W, DNS_FIELD_OMEGA};
    for (int c = 0; ok && c < 3; ++c) {
        dns_status st = dns_fetch_field(h, comps[c], plane.data());
        if (st != DNS_OK) {
            std::fclose(fp);
            return st;
        }
        ok = write_all(fp, plane.data(), sizeof(double) * plane.size());
    }
    std::vector<double> spec(2 * (size_t)om_count);
    for (int which = 0; ok && which < 2; ++which) {
        dns_status st = which == 0 ? dns_get_om2(h, spec.data()) : dns_get_fnm1(h, spec.data());
        if (st != DNS_OK) {
            std::fclose(fp);
            return st;
        }
        ok = write_all(fp, spec.data(), sizeof(double) * spec.size());
    }
    if (std::fclose(fp) != 0) ok = false;
    if (!ok) return fail(h, DNS_E_IO, "write to %s failed", path);
    return DNS_OK;
}

extern "C" dns_status dns_load_restart(dns_handle *h, const char *path)
{
    if (!h || !path) return DNS_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return fail(h, DNS_E_IO, "cannot open %s: %s", path, std::strerror(errno));
    struct Closer {
        FILE *f;
        ~Closer() { std::fclose(f); }
    } closer{fp};
    HeadInts hi{};
    if (!read_all(fp, &hi, sizeof hi)) return fail(h, DNS_E_FORMAT, "%s: truncated header", path);
    if (std::strncmp(hi.magic, "CTRB_RESTART", 12) != 0) return fail(h, DNS_E_FORMAT, "%s: bad magic", path);
    if (hi.version < 1 || hi.version > 2) return fail(h, DNS_E_FORMAT, "%s: unsupported version %u", path, hi.version);
    const size_t rb = hi.real_bytes;
    if (!((rb == 4 && hi.cplx_bytes == 8) || (rb == 8 && hi.cplx_bytes == 16)))
        return fail(h, DNS_E_FORMAT, "%s: unsupported real/cplx sizes %u/%u", path, hi.real_bytes, hi.cplx_bytes);
    const size_t want_hb = hi.version == 2 ? v2_bytes(rb) : v1_bytes(rb);
    if (hi.header_bytes != want_hb)
        return fail(h, DNS_E_FORMAT, "%s: header size %u, expected %zu", path, hi.header_bytes, want_hb);
    if (hi.Nbase != h->N || hi.N != h->N || hi.NX != h->N || hi.NZ != h->N || hi.NK != h->NK ||
        hi.NX_full != h->M || hi.NZ_full != h->M || hi.NK_full != h->NK)
        return fail(h, DNS_E_FORMAT, "%s: restart dimensions do not match this run (N=%d)", path, h->N);
    std::vector<unsigned char> rest(want_hb - sizeof hi);
    if (!read_all(fp, rest.data(), rest.size())) return fail(h, DNS_E_FORMAT, "%s: truncated header", path);
    double sc[kScalarCount];
    for (size_t i = 0; i < kScalarCount; ++i) {
        if (rb == 8) {
            std::memcpy(&sc[i], rest.data() + i * 8, 8);
        } else {
            float f;
            std::memcpy(&f, rest.data() + i * 4, 4);
            sc[i] = f;
        }
    }
    uint64_t cnt[4] = {0, 0, 0, 0};
    const size_t coff = align8(72 + kScalarCount * rb) - 72;
    std::memcpy(cnt, rest.data() + coff, (hi.version == 2 ? 4 : 3) * sizeof(uint64_t));
    const uint64_t uc_count = (uint64_t)h->NK * h->M * 3, ur_count = (uint64_t)h->M * h->M * 3;
    const uint64_t om_count = (uint64_t)h->H * h->N;
    const bool has_uc = cnt[0] != 0;
    if ((hi.version == 1 && cnt[0] != uc_count) || (hi.version == 2 && has_uc && cnt[0] != uc_count) ||
        (hi.version == 2 && cnt[3] != ur_count) || cnt[1] != om_count || cnt[2] != om_count)
        return fail(h, DNS_E_FORMAT, "%s: restart array sizes do not match this run", path);
    // UC / UR payloads: the velocity is rebuilt from om2, so they are skipped
    const uint64_t skip = (has_uc ? cnt[0] * 2 * rb : 0) + (hi.version == 2 ? cnt[3] * rb : 0);
    if (std::fseek(fp, (long)skip, SEEK_CUR) != 0) return fail(h, DNS_E_FORMAT, "%s: truncated payload", path);
    std::vector<double> om(2 * (size_t)om_count), fnv(2 * (size_t)om_count);
    for (int which = 0; which < 2; ++which) {
        std::vector<double> &dst = which == 0 ? om : fnv;
        if (rb == 8) {
            if (!read_all(fp, dst.data(), sizeof(double) * dst.size()))
                return fail(h, DNS_E_FORMAT, "%s: truncated payload", path);
        } else {
            std::vector<float> tmp(dst.size());
            if (!read_all(fp, tmp.data(), sizeof(float) * tmp.size()))
                return fail(h, DNS_E_FORMAT, "%s: truncated payload", path);
            for (size_t i = 0; i < tmp.size(); ++i) dst[i] = tmp[i];
        }
    }
    h->sc.Re = sc[0];
    h->sc.K0 = sc[1];
    h->sc.visc = sc[2];
    h->sc.t = sc[3];
    h->sc.dt = sc[4];
    h->sc.cn = sc[5];
    h->sc.cnm1 = sc[6];
    h->sc.cflnum = sc[7];
    h->sc.cflm = sc[8];
    h->sc.it = hi.it;
    h->sc.ifn = hi.ifn;
    dns_status st = dns_set_fnm1(h, fnv.data());
    if (st != DNS_OK) return st;
    return dns_set_om2(h, om.data());
}

// =====================================================================================
// measurement
// =====================================================================================
extern "C" dns_status dns_set_kernel_timing(dns_handle *h, int enabled)
{
    if (!h) return DNS_E_ARG;
    drain_timing(h);
    h->timing_every = enabled > 0 ? enabled : 0;
    for (int i = 0; i < KT_COUNT; ++i) h->seen[i] = 0;
    return DNS_OK;
}

extern "C" dns_status dns_reset_kernel_timing(dns_handle *h)
{
    if (!h) return DNS_E_ARG;
    drain_timing(h);
    for (int i = 0; i < KT_COUNT; ++i) {
        h->launches[i] = 0;
        h->total_ms[i] = 0.0;
        h->seen[i] = 0;
    }
    return DNS_OK;
}

extern "C" dns_status dns_kernel_timing(dns_handle *h, const char **names, int64_t *launches, double *total_ms,
                                        double *bytes_per_launch, int *n)
{
    if (!h || !n) return DNS_E_ARG;
    drain_timing(h);
    const double N2 = double(h->N) * double(h->N);
    // algorithmic HBM bytes per launch, with c = bytes of a stored intermediate complex
    // (16 in FP64, 8 in FP32, 4 as FP16 pairs) and 16 for the FP64 state om, fn:
    //   k_row     reads 4 fields M x H, writes 1: 5 * 0.75 N^2 * c           (60 N^2 | 30 N^2)
    //   k_col     reads G + om, fn, writes om, fn + V: 3.75 c N^2 + 32 N^2   (92 N^2 | 62 N^2)
    //   u_i u_j form (V = u, w; G = D^, C^): k_row 4 * 0.75 N^2 * c, k_col 3 c N^2 + 32 N^2
    //   k_row_cfl reads u, w in FP64: 24 N^2
    // multi-pass passes (state 16 B; c per intermediate element; M H = 0.75 N^2, M^2 = 2.25 N^2):
    //   zinv_a  om once (8 N^2) in, 4 fields out (3 c)    zinv_b  4 fields in and out (6 c)
    //   xinv_a  4 fields in (3 c), 2 planes out (4.5 c)   xinv_b  2 planes in, 1/2 plane out (5.625 c)
    //   xfwd_b  1/2 plane in, 2/3 of it out (1.875 c)     zfwd_a  kept modes in, 1 field out (1.5 c)
    //   zfwd_b  1 field in (0.75 c), om, fn in and out (32 N^2)
    //   the half-size passes 1-3 (gradient pair only, FP64): zinv_a 8 N^2 + 1.5 c, zinv_b 3 c,
    //   xinv_a 2 fields in (1.5 c), 1 plane out (2.25 c)
    const double c = h->engine == 3 ? 4.0 : h->fp32 ? 8.0 : 16.0;
    const double fr = h->div ? 3.0 : 3.75;  // the row and column kernels' intermediates, in c N^2
    // the multi-pass u_i u_j form: pass 4 reads one plane and writes two half planes (4.5 c),
    // pass 5 carries two half planes (3.75 c), pass 7 reads two fields (1.5 c + 32)
    const bool ui = h->big_uiuj;
    const double bytes[KT_COUNT] = {fr * c * N2,          (fr * c + 32.0) * N2,   24.0 * N2,
                                    (8.0 + 3.0 * c) * N2, 6.0 * c * N2,           7.5 * c * N2,
                                    (ui ? 4.5 : 5.625) * c * N2, (ui ? 3.75 : 1.875) * c * N2, 1.5 * c * N2,
                                    ((ui ? 1.5 : 0.75) * c + 32.0) * N2, (8.0 + 1.5 * c) * N2, 3.0 * c * N2,
                                    3.75 * c * N2};
    const int cnt = std::min(*n, (int)KT_COUNT);
    for (int i = 0; i < cnt; ++i) {
        if (names) names[i] = kKernelNames[i];
        if (launches) launches[i] = h->launches[i];
        if (total_ms) total_ms[i] = h->total_ms[i];
        if (bytes_per_launch) bytes_per_launch[i] = bytes[i];
    }
    *n = KT_COUNT;
    return DNS_OK;
}

extern "C" dns_status dns_synchronize(dns_handle *h)
{
    if (!h) return DNS_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DNS_OK;
}

extern "C" size_t dns_device_bytes(const dns_handle *h) { return h ? h->bytes : 0; }

// =====================================================================================
// closure (include/dns_closure.h)
// =====================================================================================
extern "C" dns_status dns_set_closure(dns_handle *h, dns_cnn *c)
{
    if (!h) return DNS_E_ARG;
    if (!c) {
        h->closure = nullptr;
        return DNS_OK;
    }
    if (h->big || h->fp32) return fail(h, DNS_E_UNSUPPORTED, "the closure runs on the FP64 fused path only");
    if (!closure_grid_supported(h->N)) return fail(h, DNS_E_UNSUPPORTED, "no closure transforms for N=%d", h->N);
    if (cnn_device(c) != h->device) return fail(h, DNS_E_ARG, "network and solver are on different devices");
    HIPCHK(h, hipSetDevice(h->device));
    const int N = h->N, H = h->H;
    if (!h->twN) {
        dns_status st;
        if ((st = dev_alloc(h, (void **)&h->twN, sizeof(cd) * N, "closure twiddles")) != DNS_OK) return st;
        if ((st = dev_alloc(h, (void **)&h->caz, sizeof(cd) * 3 * (size_t)N * H, "closure z-lines")) != DNS_OK)
            return st;
        if ((st = dev_alloc(h, (void **)&h->cbz, sizeof(cd) * (size_t)N * H, "closure rows")) != DNS_OK) return st;
        if ((st = dev_alloc(h, (void **)&h->fext, sizeof(cd) * (size_t)N * H, "closure forcing")) != DNS_OK)
            return st;
        if ((st = dev_alloc(h, (void **)&h->cfields, sizeof(float) * 3 * (size_t)N * N, "closure fields")) != DNS_OK)
            return st;
        if ((st = dev_alloc(h, (void **)&h->chp, sizeof(float) * 3 * (size_t)N * N, "closure small scales")) != DNS_OK)
            return st;
        if ((st = dev_alloc(h, (void **)&h->cpart, sizeof(double) * closure_pack_blocks(), "closure partials")) !=
            DNS_OK)
            return st;
        if ((st = dev_alloc(h, (void **)&h->cscale, sizeof(double) * 4, "closure scales")) != DNS_OK) return st;
        std::vector<cd> tw(N);
        for (int t = 0; t < N; ++t) {
            const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)N;
            tw[t] = cd{(double)cosl(a), -(double)sinl(a)};
        }
        HIPCHK(h, hipMemcpy(h->twN, tw.data(), sizeof(cd) * N, hipMemcpyHostToDevice));
    }
    if (cnn_reserve(c, (size_t)N * N) != DNS_OK) return fail(h, DNS_E_NOMEM, "closure network: %s", cnn_error(c));
    HIPCHK(h, hipDeviceSynchronize());
    h->closure = c;
    return DNS_OK;
}

extern "C" dns_status dns_closure_forcing(dns_handle *h, double *host)
{
    if (!h || !host) return DNS_E_ARG;
    if (!h->closure) return fail(h, DNS_E_STATE, "no closure network attached");
    HIPCHK(h, hipSetDevice(h->device));
    dns_status st = closure_forcing(h);
    if (st != DNS_OK) return st;
    return download_spec(h, h->fext, host);
}

extern "C" dns_status dns_set_closure_band(dns_handle *h, double k_lo, double k_hi)
{
    if (!h) return DNS_E_ARG;
    if (!(k_lo >= 0.0) || !(k_hi >= k_lo)) return fail(h, DNS_E_ARG, "need 0 <= k_lo <= k_hi");
    h->band_lo = k_lo;
    h->band_hi = k_hi;
    return DNS_OK;
}

extern "C" dns_status dns_set_closure_limiter(dns_handle *h, int enabled, double nu_dt_max)
{
    if (!h) return DNS_E_ARG;
    if (enabled && !(nu_dt_max >= 0.0)) return fail(h, DNS_E_ARG, "nu_dt_max must be >= 0");
    h->limiter = enabled ? 1 : 0;
    h->nu_dt_max = nu_dt_max;
    return DNS_OK;
}
