// mem_hip.cpp -- command-line driver of the MI355X solver; a drop-in for the
// reference's mem_cuda (cuda/mem_cuda.cu).
//
// Same command line, environment, signals, stdout lines, output folders, CSV files
// and restart.bin as mem_cuda.  The solver is reached only through the C ABI of
// include/dns_hip.h (libdns_hip.so, loaded from the directory of this binary).
//
//   mem_hip [N Re K0 STEPS CFL UPDATE ADAPT_VISC [restart.bin] [MOV]]
//   mem_hip RESTART output_folder [MOV]
//
// Environment: PHASE_TIMING, BENCH_WARMUP_STEPS, RESTART_STEPS, MOV, MOV_DIR (as in
// mem_cuda.cu:232-243, 1680-1806), plus DNS_IC=pao|k4exp (initial condition; the
// reference's PAO-k^-3 by default), DNS_SEED (seed of the k4exp spectrum) and
// DNS_PREC=fp64|fp32|bf16x3|fp16|fp16store (FP32 transforms, FP64 state; bf16x3 / fp16 run
// the transforms' radix stages on the matrix cores, fp16store keeps the intermediates as scaled
// FP16; FP64 by default), DNS_PATH=fused|multipass|uiuj|multipass-uiuj (uiuj: the FP64
// kernels with the nonlinear term in the reference's u_i u_j form; the fused ones at N=4096)
// (kernel path where both are built; automatic by default).  MOV frames
// go to $MOV_DIR/hip_<N>_<K0>_<Re>_<CFL>_U<UPDATE>/ (MOV_DIR defaults to ./movies; the
// reference's default is a home directory of its author and its prefix "cuda_").
//
// Signals (mem_cuda.cu:77-135, 1923-2081): USR1 snapshot + continue; RTMIN snapshot +
// restart.bin + continue; USR2 pause until CONT; HUP/TERM save + exit 128+sig; INT exit
// 128+sig without saving.  Handlers only set flags, polled once per step.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <sys/stat.h>
#include <thread>

#include "adapt_visc.hpp"
#include "movie.hpp"

#include <dirent.h>
#include <vector>

#include "../../include/dns_hip.h"

namespace {

volatile std::sig_atomic_t g_exit_requested = 0;
volatile std::sig_atomic_t g_exit_signal = 0;
volatile std::sig_atomic_t g_save_requested = 0;
volatile std::sig_atomic_t g_save_restart_requested = 0;
volatile std::sig_atomic_t g_pause_requested = 0;
volatile std::sig_atomic_t g_resume_requested = 0;

void on_exit_signal(int sig)
{
    g_exit_requested = 1;
    g_exit_signal = sig;
}
void on_save(int) { g_save_requested = 1; }
void on_save_restart(int) { g_save_restart_requested = 1; }
void on_pause(int) { g_pause_requested = 1; }
void on_resume(int) { g_resume_requested = 1; }

void install_signal_handlers()
{
    std::signal(SIGHUP, on_exit_signal);
    std::signal(SIGUSR1, on_save);
    std::signal(SIGUSR2, on_pause);
    std::signal(SIGCONT, on_resume);
    std::signal(SIGRTMIN, on_save_restart);
    std::signal(SIGINT, on_exit_signal);
    std::signal(SIGTERM, on_exit_signal);
}

void ignore_signals()
{
    for (int s : {SIGHUP, SIGUSR1, SIGUSR2, SIGCONT, SIGINT, SIGTERM}) std::signal(s, SIG_IGN);
    std::signal(SIGRTMIN, SIG_IGN);
}

const char *signal_name(int sig)
{
    if (sig == SIGRTMIN) return "SIGRTMIN";
    switch (sig) {
        case SIGHUP: return "SIGHUP";
        case SIGUSR1: return "SIGUSR1";
        case SIGUSR2: return "SIGUSR2";
        case SIGCONT: return "SIGCONT";
        case SIGINT: return "SIGINT";
        case SIGTERM: return "SIGTERM";
        default: return "signal";
    }
}

bool iequals(const char *a, const char *b)
{
    if (!a || !b) return false;
    for (; *a && *b; ++a, ++b)
        if (std::toupper((unsigned char)*a) != std::toupper((unsigned char)*b)) return false;
    return *a == '\0' && *b == '\0';
}

// truthy env values: true/yes/on (any case) or a non-zero integer (mem_cuda.cu:232-243)
bool flag_enabled(const char *raw)
{
    if (!raw || !raw[0]) return false;
    if (iequals(raw, "true") || iequals(raw, "yes") || iequals(raw, "on")) return true;
    return std::atoi(raw) != 0;
}

bool parse_mov(int argc, char **argv, int first)
{
    const char *raw = std::getenv("MOV");
    for (int i = first; i < argc; ++i)
        if (iequals(argv[i], "MOV")) {
            const char *env = std::getenv("MOV");
            raw = (env && env[0]) ? env : "1";
        }
    return flag_enabled(raw);
}

// "%.0E" with "E+" shortened to "E" (mem_cuda.cu:137-149)
std::string sci_no_plus(double x)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.0E", x);
    std::string s(buf);
    const size_t p = s.find("E+");
    if (p != std::string::npos) s.replace(p, 2, "E");
    return s;
}

bool make_dir(const std::string &path)
{
    if (mkdir(path.c_str(), 0775) == 0 || errno == EEXIST) return true;
    std::fprintf(stderr, "Failed to create output folder %s: %s\n", path.c_str(), std::strerror(errno));
    return false;
}

std::string join(const std::string &folder, const char *name) { return folder + "/" + name; }

// mkdir -p
bool make_dirs(const std::string &path)
{
    for (size_t i = 1; i <= path.size(); ++i)
        if (i == path.size() || path[i] == '/') {
            const std::string p = path.substr(0, i);
            if (mkdir(p.c_str(), 0775) != 0 && errno != EEXIST) {
                std::fprintf(stderr, "Failed to create folder %s: %s\n", p.c_str(), std::strerror(errno));
                return false;
            }
        }
    return true;
}

// makeMovieFolderName (mem_cuda.cu:266-282)
std::string movie_folder(int N, double K0, double Re, double CFL, int update)
{
    const char *root = std::getenv("MOV_DIR");
    char cfl[64];
    std::snprintf(cfl, sizeof cfl, "%.3f", CFL);
    const std::string name = "hip_" + std::to_string(N) + "_" + std::to_string((int)K0) + "_" + sci_no_plus(Re) + "_" +
                             cfl + "_U" + std::to_string(update);
    return join(root && root[0] ? root : "movies", name.c_str());
}

// findLastMovieFrameIndex (mem_cuda.cu:291-332): the largest NNNNN of omega_inferno_NNNNN.png
int last_movie_frame(const std::string &folder)
{
    int last = 0;
    DIR *d = opendir(folder.c_str());
    if (!d) return 0;
    const std::string pre = "omega_inferno_", suf = ".png";
    while (dirent *e = readdir(d)) {
        const std::string n = e->d_name;
        if (n.size() <= pre.size() + suf.size() || n.compare(0, pre.size(), pre) != 0 ||
            n.compare(n.size() - suf.size(), suf.size(), suf) != 0)
            continue;
        const std::string digits = n.substr(pre.size(), n.size() - pre.size() - suf.size());
        if (!std::all_of(digits.begin(), digits.end(), [](char c) { return std::isdigit((unsigned char)c) != 0; }))
            continue;
        last = std::max(last, std::atoi(digits.c_str()));
    }
    closedir(d);
    return last;
}

std::string output_folder(int N, double K0, double Re, double CFL, int it)
{
    char cfl[64];
    std::snprintf(cfl, sizeof cfl, "%.2f", CFL);
    return "output_" + std::to_string(N) + "_" + std::to_string((int)K0) + "_" + sci_no_plus(Re) + "_" + cfl + "_" +
           std::to_string(it);
}

std::vector<std::string> split_csv(const std::string &line)
{
    std::vector<std::string> cols;
    size_t b = 0;
    while (b <= line.size()) {
        const size_t c = line.find(',', b);
        if (c == std::string::npos) {
            cols.push_back(line.substr(b));
            break;
        }
        cols.push_back(line.substr(b, c - b));
        b = c + 1;
    }
    return cols;
}

bool read_line(FILE *fp, std::string &line)
{
    char buf[4096];
    if (!std::fgets(buf, sizeof buf, fp)) return false;
    line = buf;
    while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
    return true;
}

struct RestartArgs {
    int N = 0;
    double Re = 0, K0 = 0, CFL = 0;
    int UPDATE = 100;
    std::string restart_path;
};

// RESTART folder: N, K0, Re, CFL, Update from meta.csv cols 1,2,3,4,8 (mem_cuda.cu:374-424)
bool load_restart_args(const char *folder, RestartArgs &a)
{
    if (!folder || !folder[0]) {
        std::fprintf(stderr, "RESTART requires an output folder\n");
        return false;
    }
    const std::string meta = join(folder, "meta.csv");
    FILE *fp = std::fopen(meta.c_str(), "r");
    if (!fp) {
        std::perror("loadRestartFolderArgs: fopen meta.csv");
        return false;
    }
    std::string header, row;
    const bool ok = read_line(fp, header) && read_line(fp, row);
    std::fclose(fp);
    if (!ok) {
        std::fprintf(stderr, "loadRestartFolderArgs: invalid %s\n", meta.c_str());
        return false;
    }
    const std::vector<std::string> cols = split_csv(row);
    if (cols.size() < 9) {
        std::fprintf(stderr, "loadRestartFolderArgs: expected at least 9 columns in %s\n", meta.c_str());
        return false;
    }
    a.N = std::atoi(cols[1].c_str());
    a.K0 = std::atof(cols[2].c_str());
    a.Re = std::atof(cols[3].c_str());
    a.CFL = std::atof(cols[4].c_str());
    a.UPDATE = std::atoi(cols[8].c_str());
    if (a.N <= 0 || a.K0 <= 0 || a.Re <= 0 || a.CFL <= 0 || a.UPDATE <= 0) {
        std::fprintf(stderr, "loadRestartFolderArgs: invalid run parameters in %s\n", meta.c_str());
        return false;
    }
    a.restart_path = join(folder, "restart.bin");
    FILE *rf = std::fopen(a.restart_path.c_str(), "rb");
    if (!rf) {
        std::perror("loadRestartFolderArgs: fopen restart.bin");
        return false;
    }
    std::fclose(rf);
    return true;
}

struct MetricsRow {
    int N, K0;
    double Re, CFL, VISC, DT;
    int STEPS, PALIN, SIG;
    double TIME, MINUTES, FPS;
};

struct Run {
    dns_handle *h = nullptr;
    int N = 0, M = 0, UPDATE = 100;
    double K0 = 0, Re = 0, CFL = 0;
    int start_iter = 0;
    std::vector<MetricsRow> rows;
    bool adapt = false;           // ADAPT_VISC=1
    adaptvisc::State av;          // its controller state (mem_cuda.cu:1808)
};

dns_scalars scalars(const Run &r)
{
    dns_scalars s{};
    dns_get_scalars(r.h, &s);
    return s;
}

bool check(dns_status st, const Run &r, const char *what)
{
    if (st == DNS_OK) return true;
    std::fprintf(stderr, "%s failed (status %d): %s\n", what, (int)st, dns_last_error(r.h));
    return false;
}

// appendMetricsRow (mem_cuda.cu:530-567, 996-1010): one row per iteration
// pal: the palinstrophy ratio (mem_cuda.cu:872-935) when ADAPT_VISC is on, else 0
void append_metrics(Run &r, double pal, double sigma, double seconds)
{
    const dns_scalars s = scalars(r);
    const int steps = s.it - r.start_iter;
    MetricsRow m{r.N, (int)s.K0, s.Re, s.cflnum, s.visc, s.dt, s.it, (int)(10000.0 * pal), (int)sigma, s.t,
                 seconds / 60.0, seconds > 0.0 ? steps / seconds : 0.0};
    if (!r.rows.empty() && r.rows.back().STEPS == s.it)
        r.rows.back() = m;
    else
        r.rows.push_back(m);
}

bool write_metrics(const std::vector<MetricsRow> &rows, const std::string &path)
{
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) {
        std::perror("saveMetricsCsv: fopen");
        return false;
    }
    std::fprintf(fp, "N,K0,Re,CFL,VISC,DT,STEPS,PALIN,SIG,TIME,MINUTES,FPS\n");
    for (const MetricsRow &m : rows)
        std::fprintf(fp, "%d,%d,%.17g,%.17g,%.17g,%.17g,%d,%d,%d,%.17g,%.17g,%.17g\n", m.N, m.K0, m.Re, m.CFL, m.VISC,
                     m.DT, m.STEPS, m.PALIN, m.SIG, m.TIME, m.MINUTES, m.FPS);
    std::fclose(fp);
    std::printf("[CSV] Wrote %s (metrics, %zu rows)\n", path.c_str(), rows.size());
    return true;
}

bool write_meta(const Run &r, const std::string &path, double sigma, double pal, double minutes, double fps)
{
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) {
        std::perror("saveMetaCsv: fopen");
        return false;
    }
    char ts[32];
    std::time_t now = std::time(nullptr);
    std::tm tmv{};
    localtime_r(&now, &tmv);
    std::strftime(ts, sizeof ts, "%Y-%m-%d %H:%M", &tmv);
    const dns_scalars s = scalars(r);
    std::fprintf(fp, "timestamp,N,K0,Re,CFL,visc,T,IT,Update,sigma,10K_pal_over_Zkmax2,minutes,FPS,backend\n");
    std::fprintf(fp, "%s,%d,%.9g,%.9g,%.9g,%.9g,%.9g,%d,%d,%d,%.9g,%.9g,%.9g,HIP\n", ts, r.N, s.K0, s.Re, s.cflnum,
                 s.visc, s.t, s.it, r.UPDATE, (int)sigma, 10000.0 * pal, minutes, fps);
    std::fclose(fp);
    std::printf("[CSV] Wrote %s (metadata)\n", path.c_str());
    return true;
}

// P5 PGM of an M x M plane, [lo, hi] -> [1, 255] with round-to-nearest; optional
// percentile clip from an 8192-bin histogram (dns_cuda.cu:643-792)
bool write_pgm(const std::vector<double> &f, int M, const std::string &path, bool clipped, double lo_pct,
               double hi_pct, int comp)
{
    double mn = 1.0e300, mx = -1.0e300;
    size_t finite = 0;
    for (double v : f)
        if (std::isfinite(v)) {
            mn = std::min(mn, v);
            mx = std::max(mx, v);
            ++finite;
        }
    if (finite == 0) mn = mx = 0.0;
    double lo = mn, hi = mx;
    if (clipped && finite > 0 && mx - mn > 1.0e-12 && lo_pct < hi_pct) {
        const int bins = 8192;
        std::vector<size_t> hist(bins, 0);
        const double inv = double(bins - 1) / (mx - mn);
        for (double v : f)
            if (std::isfinite(v)) {
                int b = int((v - mn) * inv);
                hist[std::min(std::max(b, 0), bins - 1)]++;
            }
        const size_t lo_rank = (size_t)std::floor(lo_pct * double(finite - 1));
        const size_t hi_rank = (size_t)std::floor(hi_pct * double(finite - 1));
        int lo_bin = 0, hi_bin = bins - 1;
        size_t acc = 0;
        for (int b = 0; b < bins; ++b)
            if ((acc += hist[b]) > lo_rank) {
                lo_bin = b;
                break;
            }
        acc = 0;
        for (int b = 0; b < bins; ++b)
            if ((acc += hist[b]) > hi_rank) {
                hi_bin = b;
                break;
            }
        const double w = (mx - mn) / bins;
        lo = mn + w * lo_bin;
        hi = mn + w * (hi_bin + 1);
        if (!(hi > lo)) {
            lo = mn;
            hi = mx;
        }
    }
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) {
        std::perror("fopen");
        return false;
    }
    std::fprintf(fp, "P5\n%d %d\n255\n", M, M);
    std::vector<unsigned char> px(f.size());
    const double range = hi - lo;
    for (size_t i = 0; i < f.size(); ++i) {
        if (range <= 1.0e-12) {
            px[i] = 128;
            continue;
        }
        const double norm = std::isfinite(f[i]) ? (f[i] - lo) / range : 0.5;
        int p = int(1.0 + norm * 254.0 + 0.5);
        px[i] = (unsigned char)std::min(std::max(p, 1), 255);
    }
    std::fwrite(px.data(), 1, px.size(), fp);
    std::fclose(fp);
    if (clipped)
        std::printf("[DUMP] Wrote %s (PGM, %dx%d, comp=%d, min=%g, max=%g, clip=%g..%g)\n", path.c_str(), M, M, comp,
                    mn, mx, lo, hi);
    else
        std::printf("[DUMP] Wrote %s (PGM, %dx%d, comp=%d, min=%g, max=%g)\n", path.c_str(), M, M, comp, mn, mx);
    return true;
}

// energy_spectrum.csv (mem_cuda.cu:575-649): shell sums of |u^|^2 + |w^|^2 of the
// M x M R2C of the physical velocity.  That transform equals M^2 times the retained
// velocity spectrum (exact zeros elsewhere), so it is formed from om2 directly; the
// mode counts still run over the whole M x (M/2 + 1) grid.
bool write_spectrum(Run &r, const std::string &path)
{
    const int N = r.N, H = N / 2, M = r.M, NK = M / 2 + 1;
    std::vector<double> om(2 * (size_t)N * H);
    if (!check(dns_get_om2(r.h, om.data()), r, "dns_get_om2")) return false;
    const int nbins = std::max(32, 2 * M);
    const double rmax = std::sqrt(2.0), knyq = 0.5 * M, M4 = double(M) * M * double(M) * M;
    std::vector<double> esum(nbins, 0.0);
    std::vector<int> cnt(nbins, 0);
    auto bin_of = [&](int kx, int kz) {
        const double rr = std::sqrt(double(kx) * kx + double(kz) * kz) / knyq;
        int b = (int)std::floor(rr / rmax * nbins);
        return std::min(std::max(b, 0), nbins - 1);
    };
    for (int z = 0; z < M; ++z) {
        const int kz = (z <= M / 2) ? z : z - M;
        for (int kx = 0; kx < NK; ++kx) {
            if (kx == 0 && kz == 0) continue;
            cnt[bin_of(kx, kz)] += (kx > 0 && kx < M / 2) ? 2 : 1;
        }
    }
    for (int iz = 0; iz < N; ++iz) {
        const int kz = (iz < H) ? iz : iz - N;
        for (int ix = 0; ix < H; ++ix) {
            if (ix == 0 && kz == 0) continue;
            const double ore = om[2 * ((size_t)iz * H + ix)], oim = om[2 * ((size_t)iz * H + ix) + 1];
            double p;
            if (ix >= 1) {
                p = (ore * ore + oim * oim) / (double(ix) * ix + double(kz) * kz);  // |u|^2 + |w|^2 = |om|^2 / K^2
            } else {
                p = (kz != 0) ? (ore * ore + oim * oim) / (double(kz) * kz) : 0.0;
            }
            p *= M4;
            if (ix > 0) p *= 2.0;
            esum[bin_of(ix, kz)] += p;
        }
    }
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) {
        std::perror("saveEnergySpectrumUvCsv: fopen");
        return false;
    }
    std::fprintf(fp, "normalized_radius,shell_sum_energy,count\n");
    for (int i = 0; i < nbins; ++i)
        if (cnt[i] > 0 && esum[i] > 0.0) {
            const double r0 = rmax * i / nbins, r1 = rmax * (i + 1) / nbins;
            std::fprintf(fp, "%.17g,%.17g,%d\n", 0.5 * (r0 + r1), esum[i], cnt[i]);
        }
    std::fclose(fp);
    std::printf("[CSV] Wrote %s (energy spectrum, %d bins)\n", path.c_str(), nbins);
    return true;
}

// dumpRunOutputsToFolder (mem_cuda.cu:1396-1435)
bool dump_outputs(Run &r, const std::string &folder, double seconds, double fps)
{
    if (!make_dir(folder)) return false;
    std::printf("[SAVE] Dumping fields to folder: %s\n", folder.c_str());
    std::vector<double> plane((size_t)r.M * r.M);
    struct Item {
        int field;
        const char *name;
        bool clip;
        int comp;
    };
    const Item items[] = {{DNS_FIELD_U, "u_velocity.pgm", false, 0},
                          {DNS_FIELD_W, "v_velocity.pgm", false, 1},
                          {DNS_FIELD_KINETIC, "kinetic.pgm", false, 2},
                          {DNS_FIELD_OMEGA, "omega.pgm", true, 2},
                          {DNS_FIELD_PSI, "stream.pgm", false, 2}};
    for (const Item &it : items) {
        if (!check(dns_fetch_field(r.h, it.field, plane.data()), r, "dns_fetch_field")) return false;
        if (!write_pgm(plane, r.M, join(folder, it.name), it.clip, 0.001, 0.999, it.comp)) return false;
    }
    // sigma and the palinstrophy ratio of the last metrics row, or computed when there is
    // none (mem_cuda.cu:1419-1422)
    double sigma = 0.0, pal = 0.0;
    if (r.rows.empty()) {
        if (!check(dns_display_sigma(r.h, &sigma), r, "dns_display_sigma")) return false;
        if (!check(dns_pal_over_ens_kmax2(r.h, &pal), r, "dns_pal_over_ens_kmax2")) return false;
    } else {
        sigma = r.rows.back().SIG;
        pal = r.rows.back().PALIN / 10000.0;
    }
    return write_spectrum(r, join(folder, "energy_spectrum.csv")) && write_metrics(r.rows, join(folder, "metrics.csv")) &&
           write_meta(r, join(folder, "meta.csv"), sigma, pal, seconds / 60.0, fps);
}

bool snapshot(Run &r, bool with_restart, double seconds, double fps)
{
    const dns_scalars s = scalars(r);
    std::string folder = output_folder(r.N, r.K0, r.Re, r.CFL, s.it);
    if (with_restart) {
        std::printf("[SIGNAL] Received SIGRTMIN; saving images and restart binary, then continuing.\n");
        folder += "_restart";
        if (!make_dir(folder)) return false;
        const std::string rp = join(folder, "restart.bin");
        if (!check(dns_save_restart(r.h, rp.c_str()), r, "dns_save_restart")) return false;
        std::printf("[SAVE] Wrote %s (restart binary)\n", rp.c_str());
    } else {
        std::printf("[SIGNAL] Received SIGUSR1; saving current state and continuing.\n");
    }
    double sigma = 0.0, pal = 0.0;
    if (!check(dns_display_sigma(r.h, &sigma), r, "dns_display_sigma")) return false;
    if (r.adapt && !check(dns_pal_over_ens_kmax2(r.h, &pal), r, "dns_pal_over_ens_kmax2")) return false;
    append_metrics(r, pal, sigma, seconds);
    return dump_outputs(r, folder, seconds, fps);
}

void print_kernel_timing(dns_handle *h)
{
    const char *names[8];
    int64_t launches[8];
    double ms[8], bytes[8];
    int n = 8;
    if (dns_kernel_timing(h, names, launches, ms, bytes, &n) != DNS_OK) return;
    std::printf("[PHASE] kernel,launches,total_ms,avg_us,algorithmic_GBps\n");
    for (int i = 0; i < n; ++i) {
        if (launches[i] == 0) continue;
        const double avg = ms[i] / launches[i];
        std::printf("[PHASE] %s,%lld,%.3f,%.3f,%.1f\n", names[i], (long long)launches[i], ms[i], 1e3 * avg,
                    bytes[i] / (avg * 1e-3) / 1e9);
    }
}

}  // namespace

int main(int argc, char **argv)
{
    install_signal_handlers();
    int N = 512, STEPS = 1001, UPDATE = 100;
    double Re = 10000.0, K0 = 10.0, CFL = 0.25;
    bool ADAPT_VISC = false, MOV = false;
    std::string restart_path;

    if (argc > 1 && std::strcmp(argv[1], "RESTART") == 0) {
        if (argc < 3) {
            std::fprintf(stderr, "Usage: %s RESTART output_folder\n", argv[0]);
            return 2;
        }
        RestartArgs ra;
        if (!load_restart_args(argv[2], ra)) return 1;
        N = ra.N;
        Re = ra.Re;
        K0 = ra.K0;
        CFL = ra.CFL;
        UPDATE = ra.UPDATE;
        STEPS = 1000000;
        restart_path = ra.restart_path;
        const char *es = std::getenv("RESTART_STEPS");
        if (es && es[0]) STEPS = std::atoi(es);
        MOV = parse_mov(argc, argv, 3);
    } else {
        if (argc > 1) N = std::atoi(argv[1]);
        if (argc > 2) Re = std::atof(argv[2]);
        if (argc > 3) K0 = std::atof(argv[3]);
        if (argc > 4) STEPS = std::atoi(argv[4]);
        if (argc > 5) CFL = std::atof(argv[5]);
        if (argc > 6) UPDATE = std::atoi(argv[6]);
        if (argc > 7) ADAPT_VISC = std::atoi(argv[7]) != 0;
        int mov_begin = 8;
        if (argc > 8) {
            if (iequals(argv[8], "MOV")) {
                mov_begin = 8;
            } else {
                restart_path = argv[8];
                mov_begin = 9;
            }
        }
        MOV = parse_mov(argc, argv, mov_begin);
    }
    if (UPDATE <= 0) UPDATE = 100;
    if (STEPS <= 0) STEPS = 1000000;
    const bool phase_timing = flag_enabled(std::getenv("PHASE_TIMING"));

    std::printf("--- INITIALIZING FPS_HIP ---\n");
    std::printf(" N=%d, Re=%d, K0=%d, STEPS=%d, CFL=%.3f, UPDATE=%d, ADAPT_VISC=%d, MOV=%d\n", N, (int)Re, (int)K0,
                STEPS, CFL, UPDATE, ADAPT_VISC ? 1 : 0, MOV ? 1 : 0);
    if (phase_timing) std::printf("[PHASE] PHASE_TIMING enabled; kernel times are taken with HIP events.\n");
    if (!restart_path.empty()) std::printf(" restart=%s\n", restart_path.c_str());

    Run r;
    r.N = N;
    r.M = 3 * N / 2;
    r.UPDATE = UPDATE;
    r.K0 = K0;
    r.Re = Re;
    r.CFL = CFL;
    r.adapt = ADAPT_VISC;
    // DNS_PREC=fp32 (or mixed): FP32 transforms and intermediates, FP64 state, CN/AB2, dt, E/Z
    // DNS_PREC=bf16x3 / fp16: the same with the line transforms' radix stages as DFT-matrix
    // products on the matrix cores (three BF16 products, or one scaled FP16 product)
    // DNS_PREC=fp16store: FP32 transforms with the intermediates stored as scaled FP16 pairs
    const char *prec_env = std::getenv("DNS_PREC");
    const int prec = !prec_env                                                    ? DNS_PREC_FP64
                     : (iequals(prec_env, "fp32") || iequals(prec_env, "mixed")) ? DNS_PREC_FP32
                     : iequals(prec_env, "bf16x3")                               ? DNS_PREC_BF16X3
                     : iequals(prec_env, "fp16")                                 ? DNS_PREC_FP16
                     : iequals(prec_env, "fp16store")                            ? DNS_PREC_FP16_STORE
                                                                                 : DNS_PREC_FP64;
    if (prec == DNS_PREC_FP32) std::printf(" precision: FP32 transforms, FP64 state (DNS_PREC=%s)\n", prec_env);
    if (prec == DNS_PREC_BF16X3 || prec == DNS_PREC_FP16)
        std::printf(" precision: FP32 transforms with %s matrix-core butterflies, FP64 state (DNS_PREC=%s)\n",
                    prec == DNS_PREC_BF16X3 ? "3xBF16" : "FP16", prec_env);
    if (prec == DNS_PREC_FP16_STORE)
        std::printf(" precision: FP32 transforms, intermediates as scaled FP16, FP64 state (DNS_PREC=%s)\n", prec_env);
    // DNS_PATH=fused / multipass picks the kernel path where both are built (default: auto)
    const char *path_env = std::getenv("DNS_PATH");
    const int path = !path_env ? DNS_PATH_AUTO
                     : iequals(path_env, "fused") ? DNS_PATH_FUSED
                     : iequals(path_env, "multipass") ? DNS_PATH_MULTIPASS
                     : iequals(path_env, "uiuj") ? DNS_PATH_FUSED_UIUJ
                     : iequals(path_env, "multipass-uiuj") ? DNS_PATH_MULTIPASS_UIUJ
                                                      : DNS_PATH_AUTO;
    if (path != DNS_PATH_AUTO)
        std::printf(" path: %s (DNS_PATH)\n", path == DNS_PATH_FUSED           ? "fused"
                                              : path == DNS_PATH_FUSED_UIUJ     ? "fused, u_i u_j form"
                                              : path == DNS_PATH_MULTIPASS_UIUJ ? "multi-pass, u_i u_j form"
                                                                                : "multi-pass");
    dns_config cfg{N, Re, K0, CFL, prec, 0, path};
    dns_status st = dns_create(&cfg, &r.h);
    if (st != DNS_OK) {
        std::fprintf(stderr, "dns_create: %s\n", dns_last_error(nullptr));
        std::printf("dnsCudaCreate failed\n");
        return 1;
    }
    std::printf(" dns_create: GPU memory = %.2f MiB\n", dns_device_bytes(r.h) / (1024.0 * 1024.0));
    if (phase_timing) dns_set_kernel_timing(r.h, 1);

    const char *ic_env = std::getenv("DNS_IC");
    const int ic = (ic_env && iequals(ic_env, "k4exp")) ? DNS_IC_K4EXP : DNS_IC_PAO_KM3;
    const char *seed_env = std::getenv("DNS_SEED");
    const uint64_t seed = seed_env ? std::strtoull(seed_env, nullptr, 10) : 0ull;
    static char pao_log[8192];
    if (!check(dns_init_ic(r.h, ic, seed, pao_log, sizeof pao_log), r, "dns_init_ic")) {
        std::printf("dnsCudaInit failed\n");
        dns_destroy(r.h);
        return 1;
    }
    if (ic == DNS_IC_PAO_KM3) {
        std::fputs(pao_log, stdout);
        std::printf(" PAO-kminus3 initialization OK. VISC=%12.7f\n", scalars(r).visc);
    } else {
        std::printf(" k4exp initialization OK (seed %llu). VISC=%12.7f\n", (unsigned long long)seed, scalars(r).visc);
    }

    if (!restart_path.empty()) {
        if (!check(dns_load_restart(r.h, restart_path.c_str()), r, "loadRestartBin")) {
            dns_destroy(r.h);
            return 1;
        }
        const dns_scalars s = scalars(r);
        std::printf("[LOAD] Loaded restart binary %s at iteration %d\n", restart_path.c_str(), s.it);
        r.K0 = s.K0;
        r.CFL = s.cflnum;
    } else {
        if (!check(dns_start(r.h), r, "dns_start")) {
            dns_destroy(r.h);
            return 1;
        }
        const dns_scalars s = scalars(r);
        std::printf(" [NEXTDT INIT] CFLM=%12.7f  DT=%12.7f  CN=%12.7f\n", s.cflm, s.dt, s.cn);
        std::printf(" Initial DT=%12.7f  CN=%12.7f\n", s.dt, s.cn);
    }

    using clk = std::chrono::steady_clock;
    dns_synchronize(r.h);
    if (phase_timing) dns_reset_kernel_timing(r.h);
    const clk::time_point tbegin = clk::now();
    int warmup = UPDATE;
    const char *ew = std::getenv("BENCH_WARMUP_STEPS");
    if (ew && ew[0]) warmup = std::max(0, std::atoi(ew));
    r.start_iter = scalars(r).it;
    // movie frames (mem_cuda.cu:1834-1844)
    std::string mov_folder;
    int mov_index = 0;
    const int mov_f = movie::scale_f(N);
    if (MOV) {
        mov_folder = movie_folder(N, K0, Re, CFL, UPDATE);
        if (!make_dirs(mov_folder)) {
            dns_destroy(r.h);
            return 1;
        }
        mov_index = last_movie_frame(mov_folder);
        std::printf("[MOV] Saving omega_inferno_%%05d.png every %d steps to %s (F=%d, next=%05d)\n", UPDATE,
                    mov_folder.c_str(), mov_f, mov_index + 1);
    }
    bool interrupted = false, steady_started = false;
    double paused = 0.0, steady_paused0 = 0.0;
    int steady_iter0 = r.start_iter;
    clk::time_point steady_begin{};
    std::printf("[BENCH] Wall-clock FPS timer starts after initial device synchronize.\n");
    if (warmup > 0) {
        std::printf("[BENCH] Steady FPS will exclude the first %d completed steps (override with BENCH_WARMUP_STEPS).\n",
                    warmup);
    } else {
        steady_started = true;
        steady_begin = tbegin;
        std::printf("[BENCH] Steady FPS timer starts immediately (BENCH_WARMUP_STEPS=0).\n");
    }
    auto elapsed = [&](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };

    for (int it = scalars(r).it + 1; it <= STEPS; ++it) {
        if (!check(dns_step(r.h, 1), r, "dns_step")) {
            dns_destroy(r.h);
            return 1;
        }
        if (it == 1 || it % UPDATE == 0 || it == STEPS) {
            double sigma = 0.0, pal = 0.0;
            // SIG before NEXTDT: both only read om, so the printed line is the same in either
            // order, and NEXTDT last leaves the velocity rows of its CFL pass in place for the
            // next step of the multi-pass path (display sigma reuses those buffers)
            if (!check(dns_display_sigma(r.h, &sigma), r, "sigma") || !check(dns_next_dt(r.h), r, "next_dt")) {
                dns_destroy(r.h);
                return 1;
            }
            if (r.adapt) {  // mem_cuda.cu:1871-1875
                if (!check(dns_pal_over_ens_kmax2(r.h, &pal), r, "dns_pal_over_ens_kmax2")) {
                    dns_destroy(r.h);
                    return 1;
                }
                dns_scalars sv = scalars(r);
                double Re_new = 0.0, visc_new = 0.0;
                if (adaptvisc::update(r.av, r.N, sv.K0, sv.visc, pal, &Re_new, &visc_new)) {
                    sv.Re = Re_new;
                    sv.visc = visc_new;
                    if (!check(dns_set_scalars(r.h, &sv), r, "dns_set_scalars")) {
                        dns_destroy(r.h);
                        return 1;
                    }
                }
            }
            append_metrics(r, pal, sigma, elapsed(tbegin) - paused);
            const dns_scalars s = scalars(r);
            std::printf(" ITERATION %6d T=%12.10f DT=%10.8f CN=%10.8f CFLM=%.6f Re=%12.5e VISC=%12.5e PALIN=%d SIG=%d "
                        "FPS=%.3f\n",
                        it, s.t, s.dt, s.cn, s.cflm, s.Re, s.visc, r.rows.back().PALIN, r.rows.back().SIG,
                        r.rows.back().FPS);
            if (MOV && it > 0 && it % UPDATE == 0) {  // mem_cuda.cu:1899-1910: frame of omega
                std::vector<double> plane((size_t)r.M * r.M);
                if (!check(dns_fetch_field(r.h, DNS_FIELD_OMEGA, plane.data()), r, "dns_fetch_field")) {
                    dns_destroy(r.h);
                    return 1;
                }
                int fw = 0, fh = 0;
                double fsig = 0.0;
                const std::vector<unsigned char> rgb = movie::render(plane, r.M, mov_f, &fw, &fh, &fsig);
                char name[64];
                std::snprintf(name, sizeof name, "omega_inferno_%05d.png", mov_index + 1);
                const std::string path = join(mov_folder, name);
                if (!movie::write_png(path, fw, fh, rgb)) {
                    std::fprintf(stderr, "dnsCudaSaveFieldInfernoPng: failed to write %s\n", path.c_str());
                    dns_destroy(r.h);
                    return 1;
                }
                std::printf("[MOV] Wrote %s (Inferno PNG, %dx%d, F=%d, comp=2, sigma=%d)\n", path.c_str(), fw, fh, mov_f,
                            (int)fsig);
                ++mov_index;
            }
            std::fflush(stdout);
            if (!steady_started && it >= warmup) {
                dns_synchronize(r.h);
                steady_begin = clk::now();
                steady_iter0 = it;
                steady_paused0 = paused;
                steady_started = true;
                std::printf("[BENCH] Steady FPS timer starts after iteration %d.\n", it);
            }
        }
        if (g_pause_requested) {
            g_pause_requested = 0;
            g_resume_requested = 0;
            dns_synchronize(r.h);
            const clk::time_point p0 = clk::now();
            std::printf("[SIGNAL] Received SIGUSR2; pausing at iteration %d. Send SIGCONT to resume.\n", it);
            std::fflush(stdout);
            while (!g_resume_requested && !g_exit_requested) std::this_thread::sleep_for(std::chrono::milliseconds(250));
            const double pd = elapsed(p0);
            paused += pd;
            if (g_resume_requested) {
                g_resume_requested = 0;
                std::printf("[SIGNAL] Received SIGCONT; resuming after %.3f paused seconds.\n", pd);
                std::fflush(stdout);
            }
        }
        if (g_exit_requested) {
            interrupted = true;
            break;
        }
        for (int which = 0; which < 2; ++which) {
            volatile std::sig_atomic_t &flag = which == 0 ? g_save_restart_requested : g_save_requested;
            if (!flag) continue;
            flag = 0;
            dns_synchronize(r.h);
            const double secs = elapsed(tbegin) - paused;
            const int done = scalars(r).it - r.start_iter;
            const double fps = secs > 0.0 ? done / secs : 0.0;
            if (!snapshot(r, which == 0, secs, fps)) {
                dns_destroy(r.h);
                return 1;
            }
            std::fflush(stdout);
        }
    }

    dns_synchronize(r.h);
    const clk::time_point tend = clk::now();
    const double elap = std::chrono::duration<double>(tend - tbegin).count() - paused;
    const dns_scalars s = scalars(r);
    const int done = s.it - r.start_iter;
    const double fps = elap > 0.0 ? done / elap : 0.0;
    const int steady_done = s.it - steady_iter0;
    double steady_elap = 0.0, steady_fps = 0.0;
    if (steady_started && steady_done > 0) {
        steady_elap = std::chrono::duration<double>(tend - steady_begin).count() - (paused - steady_paused0);
        steady_fps = steady_elap > 0.0 ? steady_done / steady_elap : 0.0;
    }
    std::printf(" Frames per second (FPS)            = %12.7f\n", fps);
    if (steady_started && steady_done > 0)
        std::printf(" Steady frames per second (FPS)     = %12.7f  (%d steps after warmup)\n", steady_fps, steady_done);
    std::printf(" Elapsed time (s): %12.7f\n", elap);
    std::printf(" FPS: %12.7f\n", fps);
    if (steady_started && steady_done > 0) {
        std::printf(" Steady elapsed time (s): %12.7f\n", steady_elap);
        std::printf(" Steady FPS: %12.7f\n", steady_fps);
    }
    std::printf(" Final T=%12.7f  CN=%12.7f  DT=%12.7f  VISC=%12.7f\n", s.t, s.cn, s.dt, s.visc);
    if (phase_timing) print_kernel_timing(r.h);
    std::fflush(stdout);

    if (interrupted) {
        ignore_signals();
        const int sig = (int)g_exit_signal;
        if (!(sig == SIGHUP || sig == SIGTERM)) {
            std::printf("[SIGNAL] Received %s (%d); exiting without saving.\n", signal_name(sig), sig);
            dns_destroy(r.h);
            return 128 + sig;
        }
        std::printf("[SIGNAL] Received %s (%d); dumping current state before exit.\n", signal_name(sig), sig);
        double sigma = 0.0, pal = 0.0;
        if (!check(dns_next_dt(r.h), r, "next_dt") || !check(dns_display_sigma(r.h, &sigma), r, "sigma") ||
            (r.adapt && !check(dns_pal_over_ens_kmax2(r.h, &pal), r, "dns_pal_over_ens_kmax2"))) {
            dns_destroy(r.h);
            return 1;
        }
        append_metrics(r, pal, sigma, elap);
        const bool ok = dump_outputs(r, output_folder(r.N, r.K0, r.Re, r.CFL, scalars(r).it), elap, fps);
        dns_destroy(r.h);
        return ok ? 128 + sig : 1;
    }
    std::printf("[OUTPUT] Normal completion; no output files written.\n");
    dns_destroy(r.h);
    return 0;
}
