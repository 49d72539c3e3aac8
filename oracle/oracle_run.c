/*
 * oracle_run.c -- command-line front end of the FP64 CPU oracle (test
 * infrastructure only).  Used for the CPU baseline timing and for manual runs.
 *
 *   oracle_run N Re K0 STEPS CFL UPDATE [ic] [form] [seed]
 *     ic   : 0 = PAO-k^-3 (reference IC), 1 = k^4 exp config IC (default 1)
 *     form : 0 = reference uiuj form, 1 = advective form (default 1)
 * Prints one line per diagnostic step and the median wall time per step.
 */
#include "dns_oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int cmpd(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int main(int argc, char **argv)
{
    if (argc < 7) {
        fprintf(stderr, "usage: %s N Re K0 STEPS CFL UPDATE [ic] [form] [seed]\n", argv[0]);
        return 2;
    }
    int N = atoi(argv[1]);
    double Re = atof(argv[2]), K0 = atof(argv[3]);
    int steps = atoi(argv[4]);
    double cfl = atof(argv[5]);
    int update = atoi(argv[6]);
    int ic = argc > 7 ? atoi(argv[7]) : 1;
    int form = argc > 8 ? atoi(argv[8]) : 1;
    unsigned long long seed = argc > 9 ? strtoull(argv[9], NULL, 10) : 0ull;
    or_state *S = or_create(N, Re, K0, cfl);
    if (!S) {
        fprintf(stderr, "or_create failed\n");
        return 1;
    }
    if (ic == 0) {
        static char log[4096];
        or_init_pao(S, log, sizeof log);
        fputs(log, stdout);
    } else {
        or_init_config(S, seed);
    }
    or_start(S, form);
    or_scalars sc;
    or_get_scalars(S, &sc);
    printf(" [NEXTDT INIT] CFLM=%12.7f  DT=%12.7f  CN=%12.7f\n", sc.cflm, sc.dt, sc.cn);
    double *times = (double *)malloc(sizeof(double) * (steps > 0 ? steps : 1));
    int nt = 0;
    double row[7];
    for (int it = 1; it <= steps; ++it) {
        double t0 = now();
        int rows = or_run(S, it, update, form, row, 1);
        double t1 = now();
        /* or_run with steps=it runs exactly one step; the it==steps rule makes
         * every call diagnostic, so time only non-diagnostic-cadence steps */
        int diag = (it == 1 || it % update == 0 || it == steps);
        if (!diag) times[nt++] = t1 - t0;
        if (diag && rows) {
            printf(" ITERATION %6d T=%.10f DT=%.10e CN=%.10f CFLM=%.8f E=%.15e Z=%.15e\n", (int)row[0], row[1],
                   row[2], row[3], row[4], row[5], row[6]);
            fflush(stdout);
        }
    }
    if (nt) {
        qsort(times, nt, sizeof(double), cmpd);
        printf(" median_step_ms=%.3f steps_timed=%d\n", 1e3 * times[nt / 2], nt);
    }
    free(times);
    or_destroy(S);
    return 0;
}
