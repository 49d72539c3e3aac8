// dns_ensemble.cpp -- host side of the batched ensemble (include/dns_ensemble.h).
//
// Members are grouped by N; each group's arrays hold its members back to back and every
// hot launch (row kernel, column kernel, CFL row kernel) covers the whole group through
// the kernels' batch dimension.  Scalars live on the device (dns_ens_kernels.hip), so a
// block of pair_every steps -- snapshot, steps, training pair, NEXTDT -- is a fixed
// sequence of launches: it is captured once into a HIP graph and replayed.  The first
// block of a run differs (NEXTDT also after step 1, as in the single-run loop) and is
// issued directly.  With use_graph = 2 a whole dns_ens_advance call -- every block of it,
// the first one included -- is one graph, launched once.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dns_ensemble.h"
#include "dns_ens_internal.h"
#include "dns_ens_kernels.h"
#include "dns_kernels.h"

using dnsfft::cd;

namespace {

thread_local std::string g_ens_create_error;

struct Group {
    int N = 0, H = 0, M = 0, n = 0, Hc = 0, Mc = 0, count = 0;
    std::vector<int> members;  // member index of each slot
    cd *om = nullptr, *fn = nullptr, *V = nullptr, *G = nullptr, *snap = nullptr, *tw = nullptr;
    double *cflrow = nullptr, *coef = nullptr;
    EnsScalars *sc = nullptr;
    cd *om_c = nullptr, *fn_c = nullptr, *V_c = nullptr, *G_c = nullptr, *tw_c = nullptr;
    float *pairs = nullptr;
    int variant = 0;          // launch configuration of the row / column kernels
    double tune_ms[16] = {};  // ms per step of each launch configuration (0: not timed)
};

}  // namespace

struct dns_ens {
    int device = 0;
    hipStream_t stream = nullptr;
    int pair_every = 20, ratio = 8, max_pairs = 0, use_graph = 1, autotune = 0;
    std::vector<dns_ens_member> members;
    std::vector<Group> groups;
    std::vector<std::pair<int, int>> where;  // member -> (group, slot)
    cd *staging = nullptr;
    int *slot_d = nullptr;
    int pairs_emitted = 0;
    long long steps_done = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // use_graph == 2: one graph for a whole dns_ens_advance call (run_blocks blocks, the
    // first one starting at it = 0 when run_first), launched once
    hipGraph_t run_graph = nullptr;
    hipGraphExec_t run_exec = nullptr;
    int run_blocks = 0;
    bool run_first = false;
    size_t bytes = 0;
    std::vector<void *> allocs;
    std::string err;
};

namespace {

dns_status efail(dns_ens *e, dns_status st, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (e)
        e->err = buf;
    else
        g_ens_create_error = buf;
    return st;
}

template <class T>
dns_status ealloc(dns_ens *e, T **p, size_t count, const char *what)
{
    void *q = nullptr;
    const size_t bytes = sizeof(T) * (count ? count : 1);
    hipError_t x = hipMalloc(&q, bytes);
    if (x != hipSuccess) {
        (void)hipGetLastError();
        return efail(e, DNS_E_NOMEM, "hipMalloc(%s, %.3f GiB): %s", what, bytes / 1073741824.0, hipGetErrorString(x));
    }
    e->allocs.push_back(q);
    e->bytes += bytes;
    *p = static_cast<T *>(q);
    return DNS_OK;
}

// W_M^t = e^{-2 pi i t / M} in long double, rounded once (as dns_create)
std::vector<cd> twiddles(int M)
{
    std::vector<cd> tw(M);
    for (int t = 0; t < M; ++t) {
        const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)M;
        tw[t] = cd{(double)cosl(a), -(double)sinl(a)};
    }
    return tw;
}

BatchArgs batch_of(const Group &g, const double *coef)
{
    BatchArgs b;
    b.count = g.count;
    b.s_om = (size_t)g.N * g.H;
    b.s_v = (size_t)4 * g.M * g.H;
    b.s_g = (size_t)g.M * g.H;
    b.s_cfl = (size_t)g.M;
    b.coef = coef;
    return b;
}

BatchArgs coarse_batch_of(const Group &g)
{
    BatchArgs b;
    b.count = g.count;
    b.s_om = (size_t)g.n * g.Hc;
    b.s_v = (size_t)4 * g.Mc * g.Hc;
    b.s_g = (size_t)g.Mc * g.Hc;
    b.s_cfl = (size_t)g.Mc;
    return b;
}

const int kF4[4] = {DNS_F_U, DNS_F_W, DNS_F_OX, DNS_F_OZ};

LaunchReq col_req(dns_ens *e, int H, int M, cd *om, cd *fn, const cd *G, cd *V, const cd *tw, int forward,
                  BatchArgs bt)
{
    LaunchReq c{};
    c.what = LAUNCH_COL;
    c.stream = e->stream;
    c.fp32 = 0;
    c.om = om;
    c.fn = fn;
    c.G = G;
    c.V = V;
    c.tw = tw;
    c.twT = tw;
    c.col.do_forward = forward;
    c.col.nfields = 4;
    for (int i = 0; i < 4; ++i) c.col.fields[i] = kF4[i];
    c.col.vstride = (size_t)M * H;
    c.col.divxz = 1.0 / (double(M) * double(M));
    c.batch = bt;
    return c;
}

LaunchReq row_req(dns_ens *e, int N, int H, int M, const cd *V, cd *G, const cd *tw, double *cflrow, int what,
                  BatchArgs bt)
{
    const double PI = 3.14159265358979323846;
    LaunchReq r{};
    r.what = what;
    r.stream = e->stream;
    r.fp32 = 0;
    r.V = const_cast<cd *>(V);
    r.vstride = (size_t)M * H;
    r.Gw = G;
    r.tw = tw;
    r.twT = tw;
    r.cflrow = cflrow;
    r.DX = r.DZ = 2.0 * PI / double(N);
    r.batch = bt;
    return r;
}

dns_status klaunch(dns_ens *e, int H, const LaunchReq &r)
{
    hipError_t x = dns_kernels_launch(H, r);
    if (x != hipSuccess) return efail(e, DNS_E_HIP, "ensemble kernel launch (%d): %s", r.what, hipGetErrorString(x));
    return DNS_OK;
}

#define ESTEP(expr)                          \
    do {                                     \
        dns_status st_ = (expr);             \
        if (st_ != DNS_OK) return st_;       \
    } while (0)
#define EHIP(expr)                                                                              \
    do {                                                                                        \
        hipError_t x_ = (expr);                                                                 \
        if (x_ != hipSuccess) return efail(e, DNS_E_HIP, "%s: %s", #expr, hipGetErrorString(x_)); \
    } while (0)

// one step of every member of group g (the single-run dns_step body)
dns_status issue_step(dns_ens *e, Group &g)
{
    EHIP(ens_launch_pre(g.sc, g.coef, g.count, e->stream));
    LaunchReq rr = row_req(e, g.N, g.H, g.M, g.V, g.G, g.tw, nullptr, LAUNCH_ROW, batch_of(g, nullptr));
    rr.variant = g.variant;
    ESTEP(klaunch(e, g.H, rr));
    LaunchReq cr = col_req(e, g.H, g.M, g.om, g.fn, g.G, g.V, g.tw, 1, batch_of(g, g.coef));
    cr.variant = g.variant;
    ESTEP(klaunch(e, g.H, cr));
    EHIP(ens_launch_post(g.sc, g.count, e->stream));
    return DNS_OK;
}

dns_status issue_nextdt(dns_ens *e, Group &g)
{
    ESTEP(klaunch(e, g.H, row_req(e, g.N, g.H, g.M, g.V, nullptr, g.tw, g.cflrow, LAUNCH_ROW_CFL,
                                  batch_of(g, nullptr))));
    EHIP(ens_launch_nextdt(g.sc, g.cflrow, g.M, g.count, e->stream));
    return DNS_OK;
}

// training pair of the snapshot state: coarse nonlinear term of F(omega) by the n-grid
// kernels, then the 4 channels on the n x n grid
dns_status issue_pairs(dns_ens *e, Group &g)
{
    EHIP(ens_launch_truncate(g.snap, g.N, g.om_c, g.fn_c, g.n, g.count, e->stream));
    const BatchArgs cb = coarse_batch_of(g);
    ESTEP(klaunch(e, g.Hc, col_req(e, g.Hc, g.Mc, g.om_c, g.fn_c, nullptr, g.V_c, g.tw_c, 0, cb)));
    ESTEP(klaunch(e, g.Hc, row_req(e, g.n, g.Hc, g.Mc, g.V_c, g.G_c, g.tw_c, nullptr, LAUNCH_ROW, cb)));
    // forward pass with VT = c2 = c3 = 0: fn_c <- FN of F(omega); om_c is scratch
    ESTEP(klaunch(e, g.Hc, col_req(e, g.Hc, g.Mc, g.om_c, g.fn_c, g.G_c, g.V_c, g.tw_c, 1, cb)));
    EHIP(ens_launch_pairs(g.om_c, g.fn, g.N, g.fn_c, g.n, g.count, g.pairs, e->slot_d, e->max_pairs, e->stream));
    return DNS_OK;
}

// one block of pair_every steps; first: the block starting at it = 0 (NEXTDT after step 1)
dns_status issue_block(dns_ens *e, bool first)
{
    for (Group &g : e->groups)
        EHIP(ens_launch_copy(g.om, g.snap, (size_t)g.count * g.N * g.H, e->stream));
    for (Group &g : e->groups) ESTEP(issue_step(e, g));
    for (Group &g : e->groups) ESTEP(issue_pairs(e, g));
    EHIP(ens_launch_bump(e->slot_d, e->stream));
    if (first)
        for (Group &g : e->groups) ESTEP(issue_nextdt(e, g));
    for (int s = 1; s < e->pair_every; ++s)
        for (Group &g : e->groups) ESTEP(issue_step(e, g));
    for (Group &g : e->groups) ESTEP(issue_nextdt(e, g));
    return DNS_OK;
}

void drop_graph(dns_ens *e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    if (e->run_exec) (void)hipGraphExecDestroy(e->run_exec);
    if (e->run_graph) (void)hipGraphDestroy(e->run_graph);
    e->exec = nullptr;
    e->graph = nullptr;
    e->run_exec = nullptr;
    e->run_graph = nullptr;
    e->run_blocks = 0;
}

// the whole advance of `blocks` blocks (the first from it = 0 when first) as one graph
dns_status capture_run(dns_ens *e, int blocks, bool first)
{
    if (e->run_exec) (void)hipGraphExecDestroy(e->run_exec);
    if (e->run_graph) (void)hipGraphDestroy(e->run_graph);
    e->run_exec = nullptr;
    e->run_graph = nullptr;
    e->run_blocks = 0;
    EHIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    dns_status st = DNS_OK;
    for (int b = 0; b < blocks && st == DNS_OK; ++b) st = issue_block(e, first && b == 0);
    hipGraph_t g = nullptr;
    hipError_t x = hipStreamEndCapture(e->stream, &g);
    if (st != DNS_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (x != hipSuccess) return efail(e, DNS_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(x));
    e->run_graph = g;
    x = hipGraphInstantiate(&e->run_exec, e->run_graph, nullptr, nullptr, 0);
    if (x != hipSuccess) {
        (void)hipGraphDestroy(e->run_graph);
        e->run_graph = nullptr;
        return efail(e, DNS_E_HIP, "hipGraphInstantiate (whole run): %s", hipGetErrorString(x));
    }
    e->run_blocks = blocks;
    e->run_first = first;
    return DNS_OK;
}

dns_status capture_block(dns_ens *e)
{
    EHIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    dns_status st = issue_block(e, false);
    hipGraph_t g = nullptr;
    hipError_t x = hipStreamEndCapture(e->stream, &g);
    if (st != DNS_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (x != hipSuccess) return efail(e, DNS_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(x));
    e->graph = g;
    x = hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0);
    if (x != hipSuccess) {
        drop_graph(e);
        return efail(e, DNS_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(x));
    }
    return DNS_OK;
}

dns_status locate(dns_ens *e, int m, Group **g, int *slot)
{
    if (m < 0 || m >= (int)e->where.size()) return efail(e, DNS_E_ARG, "member %d out of range", m);
    *g = &e->groups[e->where[m].first];
    *slot = e->where[m].second;
    return DNS_OK;
}

// z-inverse of the 4 step spectra of every member of g (after an om change)
dns_status refresh_group(dns_ens *e, Group &g)
{
    return klaunch(e, g.H, col_req(e, g.H, g.M, g.om, g.fn, nullptr, g.V, g.tw, 0, batch_of(g, nullptr)));
}

// Times the full batched step (row pass + column kernel: z-forward, CN/AB2, z-inverse) of each
// launch configuration of g's size on scratch copies of om and fn (zero CN coefficients, so
// the copies stay bounded) and keeps the fastest: every configuration first runs 3 untimed
// steps (clocks and caches settle before any timing), then 3 rounds over all configurations
// of 5 timed steps each; a configuration's time is its best round (so one that happens to
// run while the clocks ramp is not penalised).
// plans: also the configurations with the radix-16 plan, whose results agree with the
// default's to roundoff instead of bitwise.  V is rebuilt from the real om afterwards.
dns_status autotune_group(dns_ens *e, Group &g, bool plans)
{
    const int nv = dns_kernels_variants(g.H, 0);
    const size_t bytes = sizeof(cd) * (size_t)g.count * g.N * g.H;
    cd *om = nullptr, *fn = nullptr;
    EHIP(hipMalloc(&om, bytes));
    if (hipMalloc(&fn, bytes) != hipSuccess) {
        (void)hipFree(om);
        return efail(e, DNS_E_NOMEM, "autotune scratch (%zu bytes)", bytes);
    }
    hipEvent_t a = nullptr, b = nullptr;
    dns_status st = DNS_OK;
    double best = 1e300;
    auto run = [&]() -> dns_status {
        EHIP(hipMemcpyAsync(om, g.om, bytes, hipMemcpyDeviceToDevice, e->stream));
        EHIP(hipMemcpyAsync(fn, g.fn, bytes, hipMemcpyDeviceToDevice, e->stream));
        EHIP(hipEventCreate(&a));
        EHIP(hipEventCreate(&b));
        auto steps = [&](int v, int n) -> dns_status {
            LaunchReq rr = row_req(e, g.N, g.H, g.M, g.V, g.G, g.tw, nullptr, LAUNCH_ROW, batch_of(g, nullptr));
            LaunchReq cr = col_req(e, g.H, g.M, om, fn, g.G, g.V, g.tw, 1, batch_of(g, nullptr));
            rr.variant = cr.variant = v;
            for (int k = 0; k < n; ++k) {
                ESTEP(klaunch(e, g.H, rr));
                ESTEP(klaunch(e, g.H, cr));
            }
            return DNS_OK;
        };
        auto wanted = [&](int v) { return plans || dns_kernels_variant_bitwise(v) != 0; };
        for (int v = 0; v < nv; ++v) {
            g.tune_ms[v] = 0.0;
            if (wanted(v)) ESTEP(steps(v, 3));
        }
        constexpr int rounds = 3, reps = 5;
        for (int round = 0; round < rounds; ++round)
            for (int v = 0; v < nv; ++v) {
                if (!wanted(v)) continue;
                EHIP(hipEventRecord(a, e->stream));
                ESTEP(steps(v, reps));
                EHIP(hipEventRecord(b, e->stream));
                EHIP(hipEventSynchronize(b));
                float ms = 0.0f;
                EHIP(hipEventElapsedTime(&ms, a, b));
                const double t = ms / reps;
                if (round == 0 || t < g.tune_ms[v]) g.tune_ms[v] = t;
            }
        for (int v = 0; v < nv; ++v)
            if (wanted(v) && g.tune_ms[v] < best) {
                best = g.tune_ms[v];
                g.variant = v;
            }
        return DNS_OK;
    };
    st = run();
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipStreamSynchronize(e->stream);
    (void)hipFree(om);
    (void)hipFree(fn);
    if (st != DNS_OK) return st;
    return refresh_group(e, g);  // the timed column kernels rewrote V from the scratch state
}

}  // namespace

extern "C" const char *dns_ens_last_error(const dns_ens *e) { return e ? e->err.c_str() : g_ens_create_error.c_str(); }

extern "C" void dns_ens_destroy(dns_ens *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    for (void *p : e->allocs) (void)hipFree(p);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

extern "C" dns_status dns_ens_create(const dns_ens_config *cfg, dns_ens **out)
{
    if (!cfg || !out || !cfg->members) return efail(nullptr, DNS_E_ARG, "dns_ens_create: null argument");
    *out = nullptr;
    if (cfg->nmembers < 1) return efail(nullptr, DNS_E_ARG, "nmembers must be >= 1");
    if (cfg->pair_every < 1) return efail(nullptr, DNS_E_ARG, "pair_every must be >= 1");
    if (cfg->filter_ratio < 2) return efail(nullptr, DNS_E_ARG, "filter_ratio must be >= 2");
    if (cfg->max_pairs < 0) return efail(nullptr, DNS_E_ARG, "max_pairs must be >= 0");
    if (cfg->use_graph < 0 || cfg->use_graph > 2)
        return efail(nullptr, DNS_E_ARG, "use_graph %d: expected 0, 1 or 2", cfg->use_graph);
    if (cfg->autotune < 0 || cfg->autotune >= 3 + dns_kernels_variants(256, 0))
        return efail(nullptr, DNS_E_ARG, "autotune %d: expected 0, 1, 2 or 3 + configuration", cfg->autotune);
    for (int m = 0; m < cfg->nmembers; ++m) {
        const dns_ens_member &mb = cfg->members[m];
        const int N = mb.N;
        if (N < 16 || N % 4 != 0 || !dns_kernels_supported(N / 2))
            return efail(nullptr, DNS_E_UNSUPPORTED, "member %d: N=%d has no fused kernels", m, N);
        const int n = N / cfg->filter_ratio;
        if (N % cfg->filter_ratio != 0 || n % 4 != 0 || n > 64 || !dns_kernels_supported(n / 2))
            return efail(nullptr, DNS_E_UNSUPPORTED, "member %d: coarse grid N/%d = %d is not built (4..64, fused sizes)",
                         m, cfg->filter_ratio, n);
        if (!(mb.Re > 0.0) || !(mb.K0 > 0.0) || !(mb.cflnum > 0.0))
            return efail(nullptr, DNS_E_ARG, "member %d: Re, K0 and cflnum must be positive", m);
    }
    int ndev = 0;
    hipError_t x = hipGetDeviceCount(&ndev);
    if (x != hipSuccess || ndev <= 0) return efail(nullptr, DNS_E_HIP, "no HIP device available (%s)", hipGetErrorString(x));
    if (cfg->device < 0 || cfg->device >= ndev) return efail(nullptr, DNS_E_ARG, "device %d out of range", cfg->device);

    dns_ens *e = new dns_ens();
    e->device = cfg->device;
    e->pair_every = cfg->pair_every;
    e->ratio = cfg->filter_ratio;
    e->max_pairs = cfg->max_pairs;
    e->use_graph = cfg->use_graph;
    e->autotune = cfg->autotune;
    e->members.assign(cfg->members, cfg->members + cfg->nmembers);
    auto bail = [&](dns_status st) {
        g_ens_create_error = e->err;
        dns_ens_destroy(e);
        return st;
    };
    if ((x = hipSetDevice(e->device)) != hipSuccess) return bail(efail(e, DNS_E_HIP, "hipSetDevice: %s", hipGetErrorString(x)));
    if ((x = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(efail(e, DNS_E_HIP, "hipStreamCreate: %s", hipGetErrorString(x)));

    // groups by N, in order of first appearance
    e->where.resize(cfg->nmembers);
    for (int m = 0; m < cfg->nmembers; ++m) {
        const int N = cfg->members[m].N;
        int gi = -1;
        for (size_t k = 0; k < e->groups.size(); ++k)
            if (e->groups[k].N == N) gi = (int)k;
        if (gi < 0) {
            Group g;
            g.N = N;
            g.H = N / 2;
            g.M = 3 * N / 2;
            g.n = N / e->ratio;
            g.Hc = g.n / 2;
            g.Mc = 3 * g.n / 2;
            e->groups.push_back(g);
            gi = (int)e->groups.size() - 1;
        }
        Group &g = e->groups[gi];
        e->where[m] = {gi, g.count};
        g.members.push_back(m);
        g.count += 1;
    }
    size_t stage = 0;
    dns_status st;
    for (Group &g : e->groups) {
        const size_t B = g.count, nspec = (size_t)g.N * g.H, nmid = (size_t)g.M * g.H;
        const size_t cspec = (size_t)g.n * g.Hc, cmid = (size_t)g.Mc * g.Hc;
        stage = std::max(stage, nspec);
        if ((st = ealloc(e, &g.om, B * nspec, "om2")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.fn, B * nspec, "fnm1")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.snap, B * nspec, "snapshot")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.V, B * 4 * nmid, "z-inverse fields")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.G, B * nmid, "x-forward product")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.cflrow, B * g.M, "cfl rows")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.coef, B * 3, "cn coefficients")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.sc, B, "scalars")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.tw, (size_t)g.M, "twiddles")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.om_c, B * cspec, "coarse om")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.fn_c, B * cspec, "coarse fn")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.V_c, B * 4 * cmid, "coarse fields")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.G_c, B * cmid, "coarse product")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.tw_c, (size_t)g.Mc, "coarse twiddles")) != DNS_OK) return bail(st);
        if ((st = ealloc(e, &g.pairs, (size_t)e->max_pairs * B * 4 * g.n * g.n, "training pairs")) != DNS_OK)
            return bail(st);
        const std::vector<cd> tw = twiddles(g.M), twc = twiddles(g.Mc);
        if ((x = hipMemcpy(g.tw, tw.data(), sizeof(cd) * g.M, hipMemcpyHostToDevice)) != hipSuccess ||
            (x = hipMemcpy(g.tw_c, twc.data(), sizeof(cd) * g.Mc, hipMemcpyHostToDevice)) != hipSuccess)
            return bail(efail(e, DNS_E_HIP, "twiddle upload: %s", hipGetErrorString(x)));
        if ((x = hipMemset(g.om, 0, sizeof(cd) * B * nspec)) != hipSuccess ||
            (x = hipMemset(g.fn, 0, sizeof(cd) * B * nspec)) != hipSuccess)
            return bail(efail(e, DNS_E_HIP, "state clear: %s", hipGetErrorString(x)));
    }
    if ((st = ealloc(e, &e->staging, stage, "staging")) != DNS_OK) return bail(st);
    if ((st = ealloc(e, &e->slot_d, 1, "pair slot")) != DNS_OK) return bail(st);
    if ((x = hipMemset(e->slot_d, 0, sizeof(int))) != hipSuccess)
        return bail(efail(e, DNS_E_HIP, "slot clear: %s", hipGetErrorString(x)));
    *out = e;
    return DNS_OK;
}

namespace {

dns_status upload_member_om(dns_ens *e, Group &g, int slot, const double *host)
{
    const size_t nspec = (size_t)g.N * g.H;
    EHIP(hipMemcpyAsync(e->staging, host, sizeof(cd) * nspec, hipMemcpyHostToDevice, e->stream));
    EHIP(dns_launch_transpose(e->staging, g.om + slot * nspec, g.N, g.H, e->stream));
    EHIP(hipMemsetAsync(g.fn + slot * nspec, 0, sizeof(cd) * nspec, e->stream));
    EHIP(hipStreamSynchronize(e->stream));
    return DNS_OK;
}

dns_status download_member(dns_ens *e, Group &g, const cd *src, double *host)
{
    const size_t nspec = (size_t)g.N * g.H;
    EHIP(dns_launch_transpose(src, e->staging, g.H, g.N, e->stream));
    EHIP(hipMemcpyAsync(host, e->staging, sizeof(cd) * nspec, hipMemcpyDeviceToHost, e->stream));
    EHIP(hipStreamSynchronize(e->stream));
    return DNS_OK;
}

}  // namespace

extern "C" dns_status dns_ens_init(dns_ens *e)
{
    if (!e) return DNS_E_ARG;
    EHIP(hipSetDevice(e->device));
    for (Group &g : e->groups) {
        std::vector<double> om((size_t)2 * g.N * g.H);
        std::vector<EnsScalars> sc(g.count);
        for (int s = 0; s < g.count; ++s) {
            const dns_ens_member &mb = e->members[g.members[s]];
            if (dns_ic_k4exp_host(g.N, mb.K0, mb.seed, om.data()) != DNS_OK)
                return efail(e, DNS_E_ARG, "member %d: initial condition rejected", g.members[s]);
            ESTEP(upload_member_om(e, g, s, om.data()));
            // mem_cuda.cu:1755-1765 and the k^4 exp IC's visc = 1/Re
            sc[s] = EnsScalars{0, 1, 0.0, 0.0, 0.0, 1.0, 0.0, 1.0 / mb.Re, mb.cflnum, 0.0};
        }
        EHIP(hipMemcpy(g.sc, sc.data(), sizeof(EnsScalars) * g.count, hipMemcpyHostToDevice));
        ESTEP(refresh_group(e, g));
        if (e->autotune == 1 || e->autotune == 2) ESTEP(autotune_group(e, g, e->autotune == 2));
        if (e->autotune >= 3) {
            const int v = e->autotune - 3;
            if (v >= dns_kernels_variants(g.H, 0))
                return efail(e, DNS_E_ARG, "N=%d has %d launch configuration(s), %d requested", g.N,
                             dns_kernels_variants(g.H, 0), v);
            g.variant = v;
        }
        ESTEP(issue_nextdt(e, g));
    }
    EHIP(hipMemsetAsync(e->slot_d, 0, sizeof(int), e->stream));
    EHIP(hipStreamSynchronize(e->stream));
    e->pairs_emitted = 0;
    e->steps_done = 0;
    return DNS_OK;
}

extern "C" dns_status dns_ens_set_om2(dns_ens *e, int m, const double *host)
{
    if (!e || !host) return DNS_E_ARG;
    Group *g = nullptr;
    int slot = 0;
    ESTEP(locate(e, m, &g, &slot));
    EHIP(hipSetDevice(e->device));
    ESTEP(upload_member_om(e, *g, slot, host));
    ESTEP(refresh_group(e, *g));
    EHIP(hipStreamSynchronize(e->stream));
    return DNS_OK;
}

extern "C" dns_status dns_ens_advance(dns_ens *e, int nsteps)
{
    if (!e) return DNS_E_ARG;
    if (nsteps < 0 || nsteps % e->pair_every != 0)
        return efail(e, DNS_E_ARG, "nsteps=%d must be a non-negative multiple of pair_every=%d", nsteps, e->pair_every);
    EHIP(hipSetDevice(e->device));
    const int blocks = nsteps / e->pair_every;
    if (e->use_graph == 2 && blocks > 0) {
        // the whole advance in one graph launch (captured once per block count and start)
        const bool first = e->steps_done == 0;
        if (!e->run_exec || e->run_blocks != blocks || e->run_first != first) ESTEP(capture_run(e, blocks, first));
        EHIP(hipGraphLaunch(e->run_exec, e->stream));
        e->steps_done += blocks * e->pair_every;
        e->pairs_emitted += blocks;
        return DNS_OK;
    }
    for (int b = 0; b < blocks; ++b) {
        if (e->steps_done == 0) {
            ESTEP(issue_block(e, true));
        } else if (e->use_graph) {
            if (!e->exec) ESTEP(capture_block(e));
            EHIP(hipGraphLaunch(e->exec, e->stream));
        } else {
            ESTEP(issue_block(e, false));
        }
        e->steps_done += e->pair_every;
        e->pairs_emitted += 1;
    }
    return DNS_OK;
}

extern "C" dns_status dns_ens_synchronize(dns_ens *e)
{
    if (!e) return DNS_E_ARG;
    EHIP(hipSetDevice(e->device));
    EHIP(hipStreamSynchronize(e->stream));
    return DNS_OK;
}

extern "C" dns_status dns_ens_get_scalars(dns_ens *e, int m, dns_scalars *out)
{
    if (!e || !out) return DNS_E_ARG;
    Group *g = nullptr;
    int slot = 0;
    ESTEP(locate(e, m, &g, &slot));
    EHIP(hipSetDevice(e->device));
    EnsScalars s;
    EHIP(hipMemcpyAsync(&s, g->sc + slot, sizeof s, hipMemcpyDeviceToHost, e->stream));
    EHIP(hipStreamSynchronize(e->stream));
    const dns_ens_member &mb = e->members[m];
    dns_scalars o{};
    o.N = g->N;
    o.M = g->M;
    o.H = g->H;
    o.NK = 3 * g->N / 4 + 1;
    o.it = s.it;
    o.ifn = s.ifn;
    o.Re = mb.Re;
    o.K0 = mb.K0;
    o.visc = s.visc;
    o.t = s.t;
    o.dt = s.dt;
    o.cn = s.cn;
    o.cnm1 = s.cnm1;
    o.cflnum = s.cflnum;
    o.cflm = s.cflm;
    *out = o;
    return DNS_OK;
}

extern "C" dns_status dns_ens_set_scalars(dns_ens *e, int m, const dns_scalars *in)
{
    if (!e || !in) return DNS_E_ARG;
    Group *g = nullptr;
    int slot = 0;
    ESTEP(locate(e, m, &g, &slot));
    EHIP(hipSetDevice(e->device));
    EnsScalars s{in->it, in->ifn, in->t, in->dt, in->dt, in->cn, in->cnm1, in->visc, in->cflnum, in->cflm};
    EHIP(hipMemcpyAsync(g->sc + slot, &s, sizeof s, hipMemcpyHostToDevice, e->stream));
    EHIP(hipStreamSynchronize(e->stream));
    return DNS_OK;
}

extern "C" dns_status dns_ens_get_om2(dns_ens *e, int m, double *host)
{
    if (!e || !host) return DNS_E_ARG;
    Group *g = nullptr;
    int slot = 0;
    ESTEP(locate(e, m, &g, &slot));
    EHIP(hipSetDevice(e->device));
    return download_member(e, *g, g->om + (size_t)slot * g->N * g->H, host);
}

extern "C" dns_status dns_ens_get_fnm1(dns_ens *e, int m, double *host)
{
    if (!e || !host) return DNS_E_ARG;
    Group *g = nullptr;
    int slot = 0;
    ESTEP(locate(e, m, &g, &slot));
    EHIP(hipSetDevice(e->device));
    return download_member(e, *g, g->fn + (size_t)slot * g->N * g->H, host);
}

extern "C" dns_status dns_ens_pair_info(dns_ens *e, int m, int *count, int *n)
{
    if (!e) return DNS_E_ARG;
    Group *g = nullptr;
    int slot = 0;
    ESTEP(locate(e, m, &g, &slot));
    if (count) *count = std::min(e->pairs_emitted, e->max_pairs);
    if (n) *n = g->n;
    return DNS_OK;
}

extern "C" dns_status dns_ens_get_pairs(dns_ens *e, int m, int first, int count, float *host)
{
    if (!e || !host) return DNS_E_ARG;
    Group *g = nullptr;
    int slot = 0;
    ESTEP(locate(e, m, &g, &slot));
    const int have = std::min(e->pairs_emitted, e->max_pairs);
    if (first < 0 || count < 0 || first + count > have)
        return efail(e, DNS_E_ARG, "pairs %d..%d requested, %d stored", first, first + count - 1, have);
    EHIP(hipSetDevice(e->device));
    const size_t per = (size_t)4 * g->n * g->n;
    for (int p = 0; p < count; ++p)
        EHIP(hipMemcpyAsync(host + (size_t)p * per, g->pairs + ((size_t)(first + p) * g->count + slot) * per,
                            sizeof(float) * per, hipMemcpyDeviceToHost, e->stream));
    EHIP(hipStreamSynchronize(e->stream));
    return DNS_OK;
}

extern "C" dns_status dns_ens_tuning(dns_ens *e, int m, int *variant, double *ms, int *n)
{
    if (!e) return DNS_E_ARG;
    Group *g = nullptr;
    int slot = 0;
    ESTEP(locate(e, m, &g, &slot));
    const int nv = dns_kernels_variants(g->H, 0);
    if (variant) *variant = g->variant;
    if (ms && n)
        for (int i = 0; i < std::min(*n, nv); ++i) ms[i] = g->tune_ms[i];
    if (n) *n = nv;
    return DNS_OK;
}

extern "C" size_t dns_ens_device_bytes(const dns_ens *e) { return e ? e->bytes : 0; }

int dns_ens_group_count(const dns_ens *e) { return e ? (int)e->groups.size() : 0; }

int dns_ens_group_pairs(const dns_ens *e, int g, EnsPairView *out)
{
    if (!e || !out || g < 0 || g >= (int)e->groups.size()) return 1;
    const Group &gr = e->groups[g];
    out->pairs = gr.pairs;
    out->count = gr.count;
    out->stored = std::min(e->pairs_emitted, e->max_pairs);
    out->n = gr.n;
    return 0;
}
