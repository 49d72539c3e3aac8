/*
 * mock_dns.c -- a host-only stand-in for libdns_hip.so, used only by the CPU tests of
 * the mem_hip driver (argument handling, signals, output folders, CSV files, exit
 * codes).  It implements the C ABI of include/dns_hip.h with a toy state: om2 decays
 * by a fixed factor per step and the scalars follow the reference's NEXTDT rule with a
 * made-up CFLM.  It computes nothing physical and is not the oracle.
 *
 * MOCK_STEP_MS (environment) makes each step sleep, so tests can signal a running
 * driver.
 */
#include "../../include/dns_hip.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

struct dns_handle {
    dns_scalars sc;
    double *om;     /* N x N/2 complex */
    int have_ic;
    int step_ms;
    char err[256];
};

static char g_err[256];

const char *dns_last_error(const dns_handle *h) { return h ? h->err : g_err; }
int dns_supported_n(int N) { return N >= 8 && N % 4 == 0; }

dns_status dns_create(const dns_config *cfg, dns_handle **out)
{
    if (!cfg || !out || cfg->N < 8 || cfg->N % 4) {
        snprintf(g_err, sizeof g_err, "mock: bad N");
        return DNS_E_ARG;
    }
    dns_handle *h = (dns_handle *)calloc(1, sizeof *h);
    h->sc.N = cfg->N;
    h->sc.M = 3 * cfg->N / 2;
    h->sc.H = cfg->N / 2;
    h->sc.NK = 3 * cfg->N / 4 + 1;
    h->sc.Re = cfg->Re;
    h->sc.K0 = cfg->K0;
    h->sc.cflnum = cfg->cflnum;
    h->sc.cn = 1.0;
    h->sc.ifn = 1;
    h->om = (double *)calloc((size_t)cfg->N * cfg->N, sizeof(double));
    const char *ms = getenv("MOCK_STEP_MS");
    h->step_ms = ms ? atoi(ms) : 0;
    *out = h;
    return DNS_OK;
}

void dns_destroy(dns_handle *h)
{
    if (!h) return;
    free(h->om);
    free(h);
}

dns_status dns_init_ic(dns_handle *h, int kind, uint64_t seed, char *log, int logcap)
{
    const int n = h->sc.N * h->sc.N;
    for (int i = 0; i < n; ++i) h->om[i] = sin(0.1 * i + (double)seed) * (kind == DNS_IC_PAO_KM3 ? 1.0 : 0.5);
    h->sc.visc = 1.0 / h->sc.Re;
    h->sc.it = 0;
    h->sc.t = h->sc.dt = 0.0;
    h->sc.cn = 1.0;
    h->sc.cnm1 = 0.0;
    h->have_ic = 1;
    if (log && logcap > 0) snprintf(log, (size_t)logcap, " mock initial condition\n");
    return DNS_OK;
}

static double mock_cflm(const dns_handle *h) { return 100.0 + 10.0 * fabs(h->om[1]); }

dns_status dns_next_dt(dns_handle *h)
{
    const double PI = 3.14159265358979323846;
    const double c = mock_cflm(h);
    if (h->sc.it == 0 && h->sc.dt == 0.0) {
        h->sc.dt = h->sc.cflnum / c / PI;
    } else {
        h->sc.cn = 0.8 + 0.2 * (h->sc.cflnum / (c * h->sc.dt * PI));
        h->sc.dt *= h->sc.cn;
    }
    h->sc.cflm = c;
    return DNS_OK;
}

dns_status dns_start(dns_handle *h)
{
    if (!h->have_ic) return DNS_E_STATE;
    return dns_next_dt(h);
}

dns_status dns_step(dns_handle *h, int n)
{
    for (int s = 0; s < n; ++s) {
        h->sc.it += 1;
        const int m = h->sc.N * h->sc.N;
        for (int i = 0; i < m; ++i) h->om[i] *= 0.999;
        h->sc.t += h->sc.dt;
        h->sc.cnm1 = h->sc.cn;
        if (h->step_ms > 0) {
            struct timespec ts = {h->step_ms / 1000, (long)(h->step_ms % 1000) * 1000000L};
            nanosleep(&ts, NULL);
        }
    }
    return DNS_OK;
}

dns_status dns_energy_enstrophy(dns_handle *h, double *E, double *Z)
{
    double s = 0.0;
    for (int i = 0; i < h->sc.N * h->sc.N; ++i) s += h->om[i] * h->om[i];
    *E = s;
    *Z = 2.0 * s;
    return DNS_OK;
}

dns_status dns_display_sigma(dns_handle *h, double *sigma)
{
    *sigma = 20.0 + fabs(h->om[2]);
    return DNS_OK;
}

/* toy palinstrophy ratio: 10000 x ratio = 30, above the controller's target of 20 */
dns_status dns_pal_over_ens_kmax2(dns_handle *h, double *out)
{
    (void)h;
    *out = 0.003;
    return DNS_OK;
}

dns_status dns_fetch_field(dns_handle *h, int field, double *host)
{
    const int M = h->sc.M;
    for (int z = 0; z < M; ++z)
        for (int x = 0; x < M; ++x) host[(size_t)z * M + x] = sin(0.05 * x * (field + 1)) * cos(0.03 * z) * h->om[1];
    return DNS_OK;
}

dns_status dns_get_scalars(const dns_handle *h, dns_scalars *out)
{
    *out = h->sc;
    return DNS_OK;
}

dns_status dns_set_scalars(dns_handle *h, const dns_scalars *in)
{
    const int N = h->sc.N, M = h->sc.M, H = h->sc.H, NK = h->sc.NK;
    h->sc = *in;
    h->sc.N = N;
    h->sc.M = M;
    h->sc.H = H;
    h->sc.NK = NK;
    return DNS_OK;
}

dns_status dns_get_om2(dns_handle *h, double *host)
{
    memcpy(host, h->om, sizeof(double) * (size_t)h->sc.N * h->sc.N);
    return DNS_OK;
}

dns_status dns_set_om2(dns_handle *h, const double *host)
{
    memcpy(h->om, host, sizeof(double) * (size_t)h->sc.N * h->sc.N);
    h->have_ic = 1;
    return DNS_OK;
}

dns_status dns_get_fnm1(dns_handle *h, double *host)
{
    memset(host, 0, sizeof(double) * (size_t)h->sc.N * h->sc.N);
    return DNS_OK;
}

dns_status dns_set_fnm1(dns_handle *h, const double *host) { (void)h; (void)host; return DNS_OK; }

/* the mock's own restart file: "MOCKRST" + scalars + om2 */
dns_status dns_save_restart(dns_handle *h, const char *path)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return DNS_E_IO;
    fwrite("MOCKRST", 1, 8, fp);
    fwrite(&h->sc, sizeof h->sc, 1, fp);
    fwrite(h->om, sizeof(double), (size_t)h->sc.N * h->sc.N, fp);
    fclose(fp);
    return DNS_OK;
}

dns_status dns_load_restart(dns_handle *h, const char *path)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) {
        snprintf(h->err, sizeof h->err, "cannot open %s", path);
        return DNS_E_IO;
    }
    char magic[8];
    dns_scalars sc;
    int ok = fread(magic, 1, 8, fp) == 8 && memcmp(magic, "MOCKRST", 8) == 0 && fread(&sc, sizeof sc, 1, fp) == 1 &&
             sc.N == h->sc.N && fread(h->om, sizeof(double), (size_t)h->sc.N * h->sc.N, fp) == (size_t)h->sc.N * h->sc.N;
    fclose(fp);
    if (!ok) {
        snprintf(h->err, sizeof h->err, "%s: bad magic or dimensions", path);
        return DNS_E_FORMAT;
    }
    h->sc = sc;
    h->have_ic = 1;
    return DNS_OK;
}

int dns_ic_pao_host(int N, double Re, double K0, double *om2, char *log, int logcap, double *visc)
{
    (void)N; (void)K0; (void)om2; (void)log; (void)logcap;
    if (visc) *visc = 1.0 / Re;
    return 0;
}

dns_status dns_ic_k4exp_host(int N, double K0, uint64_t seed, double *om2)
{
    (void)N; (void)K0; (void)seed; (void)om2;
    return DNS_OK;
}

dns_status dns_set_kernel_timing(dns_handle *h, int enabled) { (void)h; (void)enabled; return DNS_OK; }
dns_status dns_reset_kernel_timing(dns_handle *h) { (void)h; return DNS_OK; }
dns_status dns_kernel_timing(dns_handle *h, const char **names, int64_t *launches, double *total_ms,
                             double *bytes_per_launch, int *n)
{
    (void)h;
    if (*n >= 1) {
        if (names) names[0] = "mock_step";
        if (launches) launches[0] = 1;
        if (total_ms) total_ms[0] = 1.0;
        if (bytes_per_launch) bytes_per_launch[0] = 1.0;
    }
    *n = 1;
    return DNS_OK;
}
dns_status dns_synchronize(dns_handle *h) { (void)h; return DNS_OK; }
size_t dns_device_bytes(const dns_handle *h) { (void)h; return 0; }
