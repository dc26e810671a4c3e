/*
 * autotune.c -- the library's autotuner.  It times lila_dgeqrf_w03_recursive
 * (VRTQ, FP64, hh leaf) on a device matrix of the caller's shape and picks,
 * one parameter at a time (coordinate descent from the current settings):
 *   - the T block height mt and the recursion cutoff nx (handed back to the
 *     caller, who owns the T layout and lila_param[3]);
 *   - the GEMM tile shapes (of products with and without a transposed A),
 *     the split-K occupancy target, the small-output tile thresholds, the
 *     slab height of the Cholesky-QR2 panel, the panel width nb (dgeqr3's
 *     recursion cutoff, 16..32) and the panel algorithm (kept in the
 *     library's tunables for the rest of the process);
 *   - for the plain FP64 form or, after lila_autotune_form(1), the mixed one.
 * No counterpart in the reference, whose mt, nx and block sizes are chosen
 * by hand on the command line of its test drivers.
 */
#define _POSIX_C_SOURCE 199309L
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/lila.h"
#include "lila_dev.h"
#include "lila_drivers.h"

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

typedef struct {
    int m, n, reps, mt, nx, trials;
    const double *A0;
    double *A, *T, *Q;
    double best;
    char *out;
    int left;
} tune_ctx;

static void logf_(tune_ctx *c, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

static void logf_(tune_ctx *c, const char *fmt, ...) {
    if (c->left <= 1) return;
    va_list ap;
    va_start(ap, fmt);
    int w = vsnprintf(c->out, (size_t)c->left, fmt, ap);
    va_end(ap);
    if (w > 0 && w < c->left) { c->out += w; c->left -= w; }
    else c->left = 0;
}

/* best of `reps` factorizations of the pristine A0 (copied into A first) */
static double time_qr(tune_ctx *c, int reps) {
    const int p[4] = {ld_tune()->tune_form == 1 ? 1 : 0, 1, 1, c->nx};
    double best = 1e300;
    for (int r = 0; r < reps; r++) {
        ld_lacpy('A', c->m, c->n, c->A0, c->m, c->A, c->m);
        ld_laset('A', c->mt, c->n, 0.0, 0.0, c->T, c->mt);
        ld_sync();
        const double t0 = now_s();
        dv_dgeqrf_w03_recursive(p, c->m, c->n, 0, c->mt, c->A, c->m, c->T, c->mt, c->Q, c->m);
        ld_sync();
        const double t = now_s() - t0;
        if (t < best) best = t;
    }
    return best;
}

/* try each value of one integer setting; keep the fastest */
static void sweep(tune_ctx *c, const char *name, int *slot, const int *vals, int nv) {
    int keep = *slot;
    for (int k = 0; k < nv; k++) {
        if (vals[k] == keep) continue;
        const int prev = *slot;
        *slot = vals[k];
        const double t = time_qr(c, c->reps);
        c->trials++;
        logf_(c, "; %s %d %.3f", name, vals[k], t * 1e3);
        if (t < c->best) { c->best = t; keep = vals[k]; }
        *slot = prev;
    }
    *slot = keep;
}

/* which form lila_autotune times: 0 plain FP64, 1 mixed precision
 * (lila_param[0] = 1); returns the previous choice */
int lila_autotune_form(int form) {
    if (ld_init() != 0) return -1;
    if (form != 0 && form != 1) return -1;
    const int old = ld_tune()->tune_form;
    ld_tune()->tune_form = form;
    return old;
}

int lila_autotune(int m, int n, int reps, int *mt, int *nx, char *report, int len) {
    if (ld_init() != 0) return -1;
    if (m < n || n < 1 || !mt || !nx || *mt < 1 || *nx < 1) return -1;
    enum { MT_MAX = 1024 };
    ld_tune_t *tu = ld_tune();
    const size_t bytes = (size_t)m * n * sizeof(double);
    double *A0 = (double *)ld_dev_malloc(bytes), *A = (double *)ld_dev_malloc(bytes);
    double *Q = (double *)ld_dev_malloc(bytes);
    const int tmax = *mt > MT_MAX ? *mt : MT_MAX;
    double *T = (double *)ld_dev_malloc((size_t)tmax * n * sizeof(double));
    if (!A0 || !A || !Q || !T) {
        ld_dev_free(A0); ld_dev_free(A); ld_dev_free(Q); ld_dev_free(T);
        return -1;
    }
    tune_ctx c = {m, n, reps < 1 ? 1 : reps, *mt, *nx, 0, A0, A, T, Q, 0.0, report,
                  (report && len > 0) ? len : 0};
    if (report && len > 0) report[0] = '\0';
    ld_fill_uniform(20241203ULL, m, n, A0, m, -1.0, 1.0);
    time_qr(&c, 1); /* warm-up: code objects, workspace */
    c.best = time_qr(&c, c.reps);
    logf_(&c, "start mt %d nx %d %.3f ms", c.mt, c.nx, c.best * 1e3);
    {
        static const int v[] = {256, 512, 1024};
        sweep(&c, "mt", &c.mt, v, 3);
    }
    {
        static const int v[] = {32, 64, 128, 256, 512, 1024, 2048};
        sweep(&c, "nx", &c.nx, v, 7);
    }
    {
        /* the tile shape is a pair: sweep an index into the table */
        static const int tiles[][2] = {{64, 128}, {64, 64}, {128, 64}, {128, 128}};
        int keep = -1;
        for (int k = 0; k < 4; k++)
            if (tiles[k][0] == tu->gemm_bm && tiles[k][1] == tu->gemm_bn) keep = k;
        const int bm0 = tu->gemm_bm, bn0 = tu->gemm_bn;
        for (int k = 0; k < 4; k++) {
            if (k == keep) continue;
            tu->gemm_bm = tiles[k][0]; tu->gemm_bn = tiles[k][1];
            const double t = time_qr(&c, c.reps);
            c.trials++;
            logf_(&c, "; tile %dx%d %.3f", tiles[k][0], tiles[k][1], t * 1e3);
            if (t < c.best) { c.best = t; keep = k; }
        }
        if (keep >= 0) { tu->gemm_bm = tiles[keep][0]; tu->gemm_bn = tiles[keep][1]; }
        else { tu->gemm_bm = bm0; tu->gemm_bn = bn0; }
    }
    {
        /* the tile of the products with a transposed A (W = V^T C) */
        static const int tiles[][2] = {{128, 128}, {64, 128}, {128, 64}};
        int keep = -1;
        for (int k = 0; k < 3; k++)
            if (tiles[k][0] == tu->gemm_bm_t && tiles[k][1] == tu->gemm_bn_t) keep = k;
        const int bm0 = tu->gemm_bm_t, bn0 = tu->gemm_bn_t;
        for (int k = 0; k < 3; k++) {
            if (k == keep) continue;
            tu->gemm_bm_t = tiles[k][0]; tu->gemm_bn_t = tiles[k][1];
            const double t = time_qr(&c, c.reps);
            c.trials++;
            logf_(&c, "; tile_t %dx%d %.3f", tiles[k][0], tiles[k][1], t * 1e3);
            if (t < c.best) { c.best = t; keep = k; }
        }
        if (keep >= 0) { tu->gemm_bm_t = tiles[keep][0]; tu->gemm_bn_t = tiles[keep][1]; }
        else { tu->gemm_bm_t = bm0; tu->gemm_bn_t = bn0; }
    }
    {
        static const int v[] = {256, 512, 1024};
        sweep(&c, "min_wgs", &tu->gemm_min_wgs, v, 3);
    }
    {
        static const int v[] = {128, 256, 512};
        sweep(&c, "small_tiles", &tu->gemm_small_tiles, v, 3);
    }
    {
        static const int v[] = {32, 64, 128, 256};
        sweep(&c, "small_tiles_t", &tu->gemm_small_tiles_t, v, 4);
    }
    {
        static const int v[] = {96, 128, 192, 256};
        sweep(&c, "panel_rows", &tu->panel_rows_cqr, v, 4);
    }
    {
        /* panel width nb (dgeqr3's recursion cutoff) and the panel algorithm */
        if ((tu->panel_nb < 8 || tu->panel_nb > 32) && tu->panel_nb != 64) tu->panel_nb = 32;
        static const int v[] = {32, 64, 16, 24};
        sweep(&c, "panel_nb", &tu->panel_nb, v, 4);
        static const int pm[] = {1, 0};
        sweep(&c, "panel_mode", &tu->panel_mode, pm, 2);
        /* Q built at every node (the reference's call tree) or after the VRT
         * recursion by whole T blocks */
        static const int qa[] = {0, 1};
        sweep(&c, "q_after", &tu->q_after, qa, 2);
    }
    logf_(&c, "; chose mt %d nx %d tile %dx%d tile_t %dx%d min_wgs %d small_tiles %d "
              "small_tiles_t %d panel_rows %d panel_nb %d panel_mode %d q_after %d (%s form): "
              "%.3f ms (%d trials)",
          c.mt, c.nx, tu->gemm_bm, tu->gemm_bn, tu->gemm_bm_t, tu->gemm_bn_t, tu->gemm_min_wgs,
          tu->gemm_small_tiles, tu->gemm_small_tiles_t, tu->panel_rows_cqr, tu->panel_nb,
          tu->panel_mode, tu->q_after, tu->tune_form == 1 ? "mixed" : "plain", c.best * 1e3, c.trials);
    *mt = c.mt;
    *nx = c.nx;
    ld_dev_free(A0); ld_dev_free(A); ld_dev_free(Q); ld_dev_free(T);
    ld_sync();
    return 0;
}
