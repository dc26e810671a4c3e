/*
 * api.c -- the exported C ABI (include/lila.h) of liblila_hip.so.
 *
 * Each entry point classifies its matrix pointers: device pointers are used
 * in place and the call returns after enqueueing (asynchronous, like a BLAS
 * call); host pointers are staged through device buffers covering exactly
 * the span the reference routine touches ((cols-1)*ld + rows doubles) and
 * copied back before the call returns.  The work of the computation is the
 * device-pointer driver of the same name in drivers.c.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lila.h"
#include "lila_dev.h"
#include "lila_drivers.h"

typedef struct {
    double *user;
    double *dev;
    size_t count;
    int staged;
} stage_t;

static int g_host_staged = 0;
/* a staging allocation failed in this call: the driver is skipped (no
 * kernel ever sees a NULL device pointer), the buffers already staged are
 * freed without copying back, and the call returns -2 */
static int g_stage_fail = 0;

static stage_t stage_in(double *p, int ld, int rows, int cols) {
    stage_t s = {p, p, 0, 0};
    if (!p || rows <= 0 || cols <= 0) return s;
    if (g_stage_fail) { s.dev = NULL; return s; }
    if (ld_is_device_ptr(p)) return s;
    s.count = (size_t)(cols - 1) * (size_t)ld + (size_t)rows;
    s.dev = (double *)ld_dev_malloc(s.count * sizeof(double));
    if (!s.dev) {
        fprintf(stderr, "lila_hip: cannot stage %zu bytes of a host operand on the device\n",
                s.count * sizeof(double));
        g_stage_fail = 1;
        return s;
    }
    ld_memcpy_h2d(s.dev, p, s.count * sizeof(double));
    s.staged = 1;
    g_host_staged = 1;
    return s;
}

static void stage_out(stage_t *s, int copy_back) {
    if (!s->staged) return;
    if (copy_back && !g_stage_fail) ld_memcpy_d2h(s->user, s->dev, s->count * sizeof(double));
    ld_dev_free(s->dev);
    s->staged = 0;
}

static int api_begin(void) {
    if (ld_init() != 0) {
        fprintf(stderr, "lila_hip: no usable HIP device; this library has no CPU path\n");
        return -1;
    }
    /* The drivers grow their workspace arena during a call and fold its
     * chunks into one (freeing the old ones) when the call ends, read device
     * status back mid-call and clear flags with memsets: a captured graph
     * would replay into freed memory.  So they refuse capture; the batched
     * plan (lila_dgels_batched_run) is the path to replay. */
    if (ld_stream_capturing()) {
        fprintf(stderr, "lila_hip: this entry point cannot run under hipGraph capture\n");
        return -6;
    }
    g_host_staged = 0;
    g_stage_fail = 0;
    ld_clear_failed();
    return 0;
}

/* for host-staged calls wait and report; device calls stay asynchronous */
static int api_end(void) {
    int r;
    if (g_stage_fail) {
        ld_sync();
        ld_clear_info();
        g_stage_fail = 0;
        return -2;
    }
    if (g_host_staged) {
        ld_sync();
        r = ld_get_info();
        ld_clear_info();
    } else {
        /* asynchronous call: device-side status arrives with lila_get_info */
        r = ld_get_info();
        if (r < 0) ld_clear_info(); else r = 0;
    }
    if (ld_ws_mark() == 0) ld_ws_consolidate();
    return r;
}

#define BEGIN() do { const int b_ = api_begin(); if (b_ != 0) return b_; } while (0)

/* ---- factorization drivers --------------------------------------------- */

int dgeqr3(int m, int n, double *A, int lda, double *T, int ldt) {
    BEGIN();
    stage_t a = stage_in(A, lda, m, n), t = stage_in(T, ldt, n, n);
    if (!g_stage_fail) {
        dv_dgeqr3(m, n, a.dev, lda, t.dev, ldt);
    }
    stage_out(&a, 1); stage_out(&t, 1);
    return api_end();
}

int lila_dgeqrf_w03_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                              double *T, int ldt, double *Q, int ldq, double *work, int lwork) {
    (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + n), t = stage_in(T, ldt, ldt, i + n);
    stage_t q = stage_in(Q, ldq, m, i + n);
    if (!g_stage_fail) {
        dv_dgeqrf_w03_recursive(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq);
    }
    stage_out(&a, 1); stage_out(&t, 1); stage_out(&q, 1);
    return api_end();
}

int lila_dgeqrf_w03_levelx(int *lila_param, int n_lvl, int i_lvl, int *nb_lvl, int m, int n, int i,
                           int mt, double *A, int lda, double *T, int ldt, double *Q, int ldq,
                           double *work, int lwork) {
    (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + n), t = stage_in(T, ldt, ldt, i + n);
    stage_t q = stage_in(Q, ldq, m, i + n);
    if (!g_stage_fail) {
        dv_dgeqrf_w03_levelx(lila_param, n_lvl, i_lvl, nb_lvl, m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev,
                             ldq);
    }
    stage_out(&a, 1); stage_out(&t, 1); stage_out(&q, 1);
    return api_end();
}

#define LEAF_ENTRY(NAME, CALL)                                                              \
    int NAME {                                                                              \
        (void)work; (void)lwork;                                                            \
        BEGIN();                                                                            \
        stage_t a = stage_in(A, lda, m, i + n), t = stage_in(T, ldt, ldt, i + n);           \
        stage_t q = stage_in(Q, ldq, m, i + n);                                             \
        if (!g_stage_fail) {                                                                \
            CALL;                                                                               \
        }                                                                                   \
        stage_out(&a, 1); stage_out(&t, 1); stage_out(&q, 1);                               \
        return api_end();                                                                   \
    }

LEAF_ENTRY(lila_dgeqr2_w03(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                           int ldt, double *Q, int ldq, double *work, int lwork),
           dv_dgeqr2_w03(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
LEAF_ENTRY(lila_dgeqrf_w03_mt(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                              double *T, int ldt, double *Q, int ldq, double *work, int lwork),
           dv_dgeqrf_w03_mt(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
LEAF_ENTRY(lila_dgeqrf_w03_mt_hh(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                 double *T, int ldt, double *Q, int ldq, double *work, int lwork),
           dv_dgeqrf_w03_mt_hh(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
LEAF_ENTRY(lila_dgeqrf_w03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                                double *Q, int ldq, double *work, int lwork),
           dv_dgeqrf_w03_mt_l(m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
LEAF_ENTRY(lila_dgeqrf_w03_mt_hr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                                 double *Q, int ldq, double *work, int lwork),
           dv_dgeqrf_w03_mt_hr(m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
LEAF_ENTRY(lila_dgeqr2_w03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                             double *Q, int ldq, double *work, int lwork),
           dv_dgeqr2_w03_l(m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
LEAF_ENTRY(lila_dgeqr2_w03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                             double *Q, int ldq, double *work, int lwork),
           dv_dgeqr2_w03_3(m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
LEAF_ENTRY(lila_dgeqr2_w03_hr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                              double *Q, int ldq, double *work, int lwork),
           dv_dgeqr2_w03_hr(m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))

/* ---- block reflector application / Q build / T ------------------------- */

int lila_dormqrf_w03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T,
                     int ldt, double *work, int lwork) {
    (void)work; (void)lwork;
    BEGIN();
    int cols = (i + k > j + n) ? i + k : j + n;
    stage_t a = stage_in(A, lda, m, cols), t = stage_in(T, ldt, ldt, i + k);
    if (!g_stage_fail) {
        dv_dormqrf_w03(m, n, k, i, j, mt, a.dev, lda, t.dev, ldt);
    }
    stage_out(&a, 1); stage_out(&t, 0);
    return api_end();
}

int lila_dormqrf_z03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T,
                     int ldt, double *B, int ldb, double *work, int lwork) {
    (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + k), t = stage_in(T, ldt, ldt, i + k);
    stage_t b = stage_in(B, ldb, m, j + n);
    if (!g_stage_fail) {
        dv_dormqrf_z03(m, n, k, i, j, mt, a.dev, lda, t.dev, ldt, b.dev, ldb);
    }
    stage_out(&a, 0); stage_out(&t, 0); stage_out(&b, 1);
    return api_end();
}

int lila_dormqrbz_w03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *Q,
                      int ldq, double *T, int ldt, double *work, int lwork) {
    (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + k), t = stage_in(T, ldt, ldt, i + k);
    stage_t q = stage_in(Q, ldq, m, j + n);
    if (!g_stage_fail) {
        dv_dormqrbz_w03(m, n, k, i, j, mt, a.dev, lda, q.dev, ldq, t.dev, ldt);
    }
    stage_out(&a, 0); stage_out(&t, 0); stage_out(&q, 1);
    return api_end();
}

int lila_dlarft_connect_w03(int m, int n, int i, int j, int mt, double *A, int lda, double *T,
                            int ldt) {
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + n), t = stage_in(T, ldt, ldt, i + n);
    if (!g_stage_fail) {
        dv_dlarft_connect_w03(m, n, i, j, mt, a.dev, lda, t.dev, ldt);
    }
    stage_out(&a, 0); stage_out(&t, 1);
    return api_end();
}

int lila_dlarft_w03(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                    double *tau) {
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + n), t = stage_in(T, ldt, ldt, i + n);
    stage_t ta = stage_in(tau, n, n, 1);
    if (!g_stage_fail) {
        dv_dlarft_w03(m, n, i, mt, a.dev, lda, t.dev, ldt, ta.dev);
    }
    stage_out(&a, 0); stage_out(&t, 1); stage_out(&ta, 0);
    return api_end();
}

int lila_dT2tau_w03(int m, int n, int i, int mt, double *T, int ldt, double *tau) {
    (void)m;
    BEGIN();
    stage_t t = stage_in(T, ldt, ldt, i + n), ta = stage_in(tau, n, n, 1);
    if (!g_stage_fail) {
        ld_T2tau(n, i, mt, t.dev, ldt, ta.dev);
    }
    stage_out(&t, 0); stage_out(&ta, 1);
    return api_end();
}

int lila_dV2tau_w03(int m, int n, int i, int mt, double *A, int lda, double *tau) {
    (void)mt;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + n), ta = stage_in(tau, n, n, 1);
    if (!g_stage_fail) {
        ld_V2tau(m, n, i, a.dev, lda, ta.dev);
    }
    stage_out(&a, 0); stage_out(&ta, 1);
    return api_end();
}

/* ---- VRT only (v03), Q from (V, T) (q03), append-columns, LAPACK-kernel
 * and qr2 families ------------------------------------------------------- */

#define VRT_ENTRY(NAME, CALL)                                                               \
    int NAME {                                                                              \
        (void)work; (void)lwork;                                                            \
        BEGIN();                                                                            \
        stage_t a = stage_in(A, lda, m, i + n), t = stage_in(T, ldt, ldt, i + n);           \
        if (!g_stage_fail) {                                                                \
            CALL;                                                                               \
        }                                                                                   \
        stage_out(&a, 1); stage_out(&t, 1);                                                 \
        return api_end();                                                                   \
    }

VRT_ENTRY(lila_dgeqrf_v03_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                    double *T, int ldt, double *work, int lwork),
          dv_dgeqrf_v03_recursive(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt))
VRT_ENTRY(lila_dgeqrf_v03_mt(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                             int ldt, double *work, int lwork),
          dv_dgeqrf_v03_mt(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt))
VRT_ENTRY(lila_dgeqr2_v03(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                          int ldt, double *work, int lwork),
          dv_dgeqr2_v03(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt))
VRT_ENTRY(lila_dgeqrf_v03_mt_hh(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                double *T, int ldt, double *work, int lwork),
          dv_dgeqrf_v03_mt_hh(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt))
VRT_ENTRY(lila_dgeqrf_v03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                               double *work, int lwork),
          dv_dgeqrf_v03_mt_l(m, n, i, mt, a.dev, lda, t.dev, ldt))
VRT_ENTRY(lila_dgeqr2_v03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                            double *work, int lwork),
          dv_dgeqr2_v03_3(m, n, i, mt, a.dev, lda, t.dev, ldt))
VRT_ENTRY(lila_dgeqr2_v03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                            double *work, int lwork),
          dv_dgeqr2_v03_3(m, n, i, mt, a.dev, lda, t.dev, ldt))

int lila_dgeqrf_v03_levelx(int *lila_param, int n_lvl, int i_lvl, int *nb_lvl, int m, int n, int i,
                           int mt, double *A, int lda, double *T, int ldt, double *work, int lwork) {
    (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + n), t = stage_in(T, ldt, ldt, i + n);
    if (!g_stage_fail) {
        dv_dgeqrf_v03_levelx(lila_param, n_lvl, i_lvl, nb_lvl, m, n, i, mt, a.dev, lda, t.dev, ldt);
    }
    stage_out(&a, 1); stage_out(&t, 1);
    return api_end();
}

/* Q builds read A and T and write Q */
#define Q03_ENTRY(NAME, CALL)                                                               \
    int NAME {                                                                              \
        (void)work; (void)lwork;                                                            \
        BEGIN();                                                                            \
        stage_t a = stage_in(A, lda, m, i + n), t = stage_in(T, ldt, ldt, i + n);           \
        stage_t q = stage_in(Q, ldq, m, i + n);                                             \
        if (!g_stage_fail) {                                                                \
            CALL;                                                                               \
        }                                                                                   \
        stage_out(&a, 0); stage_out(&t, 0); stage_out(&q, 1);                               \
        return api_end();                                                                   \
    }

Q03_ENTRY(lila_dgeqrf_q03_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                    double *T, int ldt, double *Q, int ldq, double *work, int lwork),
          dv_dgeqrf_q03_recursive(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
Q03_ENTRY(lila_dgeqrf_q03_mt(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                             int ldt, double *Q, int ldq, double *work, int lwork),
          dv_dgeqrf_q03_mt(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
Q03_ENTRY(lila_dgeqrf_q03_mt_hh(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                double *T, int ldt, double *Q, int ldq, double *work, int lwork),
          dv_dgeqrf_q03_mt_hh(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
Q03_ENTRY(lila_dgeqrf_q03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                               double *Q, int ldq, double *work, int lwork),
          dv_dgeqrf_q03_mt_l(m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
Q03_ENTRY(lila_dgeqr2_q03(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                          int ldt, double *Q, int ldq, double *work, int lwork),
          dv_dgeqr2_q03(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
Q03_ENTRY(lila_dgeqr2_q03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                            double *Q, int ldq, double *work, int lwork),
          dv_dgeqr2_q03_l(m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
Q03_ENTRY(lila_dgeqr2_q03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                            double *Q, int ldq, double *work, int lwork),
          dv_dgeqr2_q03_3(m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))

int lila_dgeqrf_q03_levelx(int *lila_param, int n_lvl, int i_lvl, int *nb_lvl, int m, int n, int i,
                           int mt, double *A, int lda, double *T, int ldt, double *Q, int ldq,
                           double *work, int lwork) {
    (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + n), t = stage_in(T, ldt, ldt, i + n);
    stage_t q = stage_in(Q, ldq, m, i + n);
    if (!g_stage_fail) {
        dv_dgeqrf_q03_levelx(lila_param, n_lvl, i_lvl, nb_lvl, m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev,
                             ldq);
    }
    stage_out(&a, 0); stage_out(&t, 0); stage_out(&q, 1);
    return api_end();
}

int lila_dgeqrf_w03_appendcols(int *lila_param, int m, int k, int n, int mt, double *A, int lda,
                               double *Q, int ldq, double *T, int ldt, double *work, int lwork) {
    (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, k + n), t = stage_in(T, ldt, ldt, k + n);
    stage_t q = stage_in(Q, ldq, m, k + n);
    if (!g_stage_fail) {
        dv_dgeqrf_w03_appendcols(lila_param, m, k, n, mt, a.dev, lda, q.dev, ldq, t.dev, ldt);
    }
    stage_out(&a, 1); stage_out(&t, 1); stage_out(&q, 1);
    return api_end();
}

int lila_dgeqrf_LAPACK_appendcols(int m, int k, int n, int mt, double *A, int lda, double *Q, int ldq,
                                  double *T, int ldt, double *work, int lwork) {
    (void)mt; (void)ldt; (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, k + n), t = stage_in(T, 1, 1, k + n);
    stage_t q = stage_in(Q, ldq, m, k + n);
    if (!g_stage_fail) {
        dv_dgeqrf_LAPACK_appendcols(m, k, n, a.dev, lda, q.dev, ldq, t.dev);
    }
    stage_out(&a, 1); stage_out(&t, 1); stage_out(&q, 1);
    return api_end();
}

int lila_dgeqrf_ker_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                              double *T, int ldt, double *Q, int ldq, double *work, int lwork) {
    (void)mt; (void)ldt; (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + n), t = stage_in(T, 1, 1, i + n);
    stage_t q = stage_in(Q, ldq, m, i + n);
    if (!g_stage_fail) {
        dv_dgeqrf_ker_recursive(lila_param, m, n, i, a.dev, lda, t.dev, q.dev, ldq);
    }
    stage_out(&a, 1); stage_out(&t, 1); stage_out(&q, 1);
    return api_end();
}

int lila_dormqrf_w00(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T, int ldt,
                     double *work, int lwork) {
    (void)mt; (void)T; (void)ldt; (void)work; (void)lwork;
    BEGIN();
    int cols = (j + n > i + k) ? j + n : i + k;
    stage_t a = stage_in(A, lda, m, cols);
    if (!g_stage_fail) {
        dv_dormqrf_tau_from_v(m, n, k, i, a.dev, lda, a.dev + i + (size_t)j * lda, lda);
    }
    stage_out(&a, 1);
    return api_end();
}

int lila_dormqrf_z00(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *B, int ldb,
                     double *T, int ldt, double *work, int lwork) {
    (void)mt; (void)T; (void)ldt; (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + k), b = stage_in(B, ldb, m, j + n);
    if (!g_stage_fail) {
        dv_dormqrf_tau_from_v(m, n, k, i, a.dev, lda, b.dev + i + (size_t)j * ldb, ldb);
    }
    stage_out(&a, 0); stage_out(&b, 1);
    return api_end();
}

int lila_dormqrf_z02(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T, int ldt,
                     double *B, int ldb, double *work, int lwork) {
    (void)mt; (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + k), t = stage_in(T, ldt, i + k, i + k);
    stage_t b = stage_in(B, ldb, m, j + n);
    if (!g_stage_fail) {
        dv_dormqrf_z02(m, n, k, i, j, a.dev, lda, t.dev, ldt, b.dev, ldb);
    }
    stage_out(&a, 0); stage_out(&t, 0); stage_out(&b, 1);
    return api_end();
}

LEAF_ENTRY(lila_dgeqrf_qr2(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                           int ldt, double *Q, int ldq, double *work, int lwork),
           dv_dgeqrf_qr2(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))
LEAF_ENTRY(lila_dgeqrf_qr2_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                     double *T, int ldt, double *Q, int ldq, double *work, int lwork),
           dv_dgeqrf_qr2_recursive(lila_param, m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq))

int lila_dgeqrf_qr2_recursive2(int *lila_param, int m, int n, int i, int mt, int nb1, int nb2, double *A,
                               int lda, double *T, int ldt, double *Q, int ldq, double *work,
                               int lwork) {
    (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + n), t = stage_in(T, ldt, ldt, i + n);
    stage_t q = stage_in(Q, ldq, m, i + n);
    if (!g_stage_fail) {
        dv_dgeqrf_qr2_recursive2(lila_param, m, n, i, mt, nb1, nb2, a.dev, lda, t.dev, ldt, q.dev, ldq);
    }
    stage_out(&a, 1); stage_out(&t, 1); stage_out(&q, 1);
    return api_end();
}

/* ---- Householder reconstruction ----------------------------------------- */

int lila_dorgh2_3(int m, int n, double *A, int lda, double *T, int ldt, double *Q, int ldq,
                  double *work, int lwork, double *S) {
    (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, n), t = stage_in(T, ldt, n, n), q = stage_in(Q, ldq, m, n);
    stage_t s = stage_in(S, n, n, 1);
    if (!g_stage_fail) {
        dv_dorgh2_3(m, n, a.dev, lda, t.dev, ldt, q.dev, ldq, s.dev);
    }
    stage_out(&a, 1); stage_out(&t, 1); stage_out(&q, 1); stage_out(&s, 1);
    return api_end();
}

int lila_dorgh2(int m, int n, int i, int j, int mt, double *A, int lda, double *T, int ldt,
                double *Q, int ldq, double *work, int lwork, double *S) {
    (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, j + n), t = stage_in(T, ldt, ldt, j + n);
    stage_t q = stage_in(Q, ldq, m, j + n), s = stage_in(S, j + n, j + n, 1);
    if (!g_stage_fail) {
        dv_dorgh2(m, n, i, j, mt, a.dev, lda, t.dev, ldt, q.dev, ldq, s.dev);
    }
    stage_out(&a, 1); stage_out(&t, 1); stage_out(&q, 1); stage_out(&s, 1);
    return api_end();
}

int lila_ormhr2_w03_hr(int m, int n, int i, int j, int l, int mt, double *A, int lda, double *T,
                       int ldt, double *Q, int ldq, double *work, int lwork, double *S) {
    (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, j + n), t = stage_in(T, ldt, ldt, j + n);
    stage_t q = stage_in(Q, ldq, m, j + n), s = stage_in(S, j + n, j + n, 1);
    if (!g_stage_fail) {
        dv_ormhr2_w03_hr(m, n, i, j, l, mt, a.dev, lda, t.dev, ldt, q.dev, ldq, s.dev);
    }
    stage_out(&a, 1); stage_out(&t, 1); stage_out(&q, 1); stage_out(&s, 1);
    return api_end();
}

int lila_dorghr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                int ldq, double *work, int lwork, int *S) {
    (void)work; (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + n), t = stage_in(T, ldt, ldt, i + n);
    stage_t q = stage_in(Q, ldq, m, i + n);
    if (!g_stage_fail) {
        size_t mark = ld_ws_mark();
        double *Sd = (double *)ld_ws_alloc(sizeof(double) * (size_t)(i + n));
        dv_dorghr(m, n, i, mt, a.dev, lda, t.dev, ldt, q.dev, ldq, Sd);
        if (S) {
            double *hs = (double *)malloc(sizeof(double) * (size_t)(i + n));
            ld_memcpy_d2h(hs, Sd, sizeof(double) * (size_t)(i + n));
            int dev_s = ld_is_device_ptr(S);
            int *hi = dev_s ? (int *)malloc(sizeof(int) * (size_t)(i + n)) : S;
            for (int k = i; k < i + n; k++) hi[k] = (int)hs[k];
            if (dev_s) {
                ld_memcpy_h2d(S + i, hi + i, sizeof(int) * (size_t)n);
                free(hi);
            }
            free(hs);
        }
        ld_ws_release(mark);
    }
    stage_out(&a, 1); stage_out(&t, 1); stage_out(&q, 1);
    return api_end();
}

int lila_dgetrf_b(int m, int n, int i, int mt, double *A, int lda, double *Q, int ldq, double *work,
                  int lwork) {
    (void)lwork;
    BEGIN();
    stage_t a = stage_in(A, lda, m, i + n), q = stage_in(Q, ldq, m, i + n);
    stage_t u = stage_in(work, n, n, n);
    if (!g_stage_fail) {
        dv_dgetrf_b(m, n, i, mt, a.dev, lda, q.dev, ldq, u.dev);
    }
    stage_out(&a, 1); stage_out(&q, 1); stage_out(&u, 1);
    return api_end();
}

/* ---- the recursive QR library of qr3/src (qr2.h) -------------------------- */

int qr2_dgeqr3_R(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr) {
    BEGIN();
    stage_t a = stage_in(A, lda, m, n), t = stage_in(T, ldt, n, n);
    const int alias = (R == A);
    stage_t r = alias ? a : stage_in(R, ldr, n, n);
    if (!g_stage_fail) {
        dv_qr2_dgeqr3_R(m, n, a.dev, lda, t.dev, ldt, alias ? a.dev : r.dev, alias ? lda : ldr);
    }
    stage_out(&a, 1); stage_out(&t, 1);
    if (!alias) stage_out(&r, 1);
    return api_end();
}

int qr2_dlarft3(int m, int n, double *A, int lda, double *T, int ldt, double *tau) {
    BEGIN();
    stage_t a = stage_in(A, lda, m, n), t = stage_in(T, ldt, n, n), u = stage_in(tau, n, n, 1);
    if (!g_stage_fail) {
        dv_qr2_dlarft3(m, n, a.dev, lda, t.dev, ldt, u.dev);
    }
    stage_out(&a, 0); stage_out(&t, 1); stage_out(&u, 0);
    return api_end();
}

int qr2_dorgqr3(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau) {
    BEGIN();
    stage_t q = stage_in(Q, ldq, m, n), t = stage_in(T, ldt, n, n), u = stage_in(tau, n, n, 1);
    if (!g_stage_fail) {
        dv_qr2_dorgqr3(m, n, q.dev, ldq, t.dev, ldt, u.dev);
    }
    stage_out(&q, 1); stage_out(&t, 0); stage_out(&u, 0);
    return api_end();
}

/* qr3/src/qr2_dorgqr3_VT2Q.c: the same Q (T's diagonal blocks are copied
 * into Q's upper part on the way and consumed there) */
int qr2_dorgqr3_VT2Q(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau) {
    return qr2_dorgqr3(m, n, Q, ldq, T, ldt, tau);
}

#define QR2_R_ENTRY(NAME, CALL)                                                       \
    int NAME(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr) { \
        BEGIN();                                                                       \
        stage_t a = stage_in(A, lda, m, n), t = stage_in(T, ldt, n, n);               \
        const int alias = (R == A);                                                    \
        stage_t r = alias ? a : stage_in(R, ldr, n, n);                                \
        double *rd = alias ? a.dev : r.dev;                                            \
        const int rl = alias ? lda : ldr;                                              \
        (void)rd; (void)rl;                                                            \
        if (!g_stage_fail) {                                                           \
            CALL;                                                                      \
        }                                                                              \
        stage_out(&a, 1); stage_out(&t, 1);                                            \
        if (!alias) stage_out(&r, 1);                                                  \
        return api_end();                                                              \
    }
QR2_R_ENTRY(qr2_dgeqr3_R_UT, dv_qr2_dgeqr3_R_UT(m, n, a.dev, lda, t.dev, ldt, rd, rl))
QR2_R_ENTRY(qr2_dgeqr3_R_ISW, dv_qr2_dgeqr3_R_isw(0, m, n, a.dev, lda, t.dev, ldt, rd, rl))
QR2_R_ENTRY(qr2_dgeqr3_R_UT_ISW, dv_qr2_dgeqr3_R_isw(1, m, n, a.dev, lda, t.dev, ldt, rd, rl))
#undef QR2_R_ENTRY

#define QR2_T_ENTRY(NAME, CALL)                                                          \
    int NAME(int m, int n, double *A, int lda, double *T, int ldt, double *tau) {        \
        BEGIN();                                                                          \
        stage_t a = stage_in(A, lda, m, n), t = stage_in(T, ldt, n, n),                   \
                u = stage_in(tau, n, n, 1);                                               \
        if (!g_stage_fail) {                                                              \
            CALL;                                                                             \
        }                                                                                 \
        stage_out(&a, 0); stage_out(&t, 1); stage_out(&u, 0);                             \
        return api_end();                                                                 \
    }
QR2_T_ENTRY(qr2_dlarft3_UT, dv_qr2_dlarft3_UT(m, n, a.dev, lda, t.dev, ldt, u.dev))
QR2_T_ENTRY(qr2_dlarft3_ISW, dv_qr2_dlarft3_isw(0, m, n, a.dev, lda, t.dev, ldt, u.dev))
QR2_T_ENTRY(qr2_dlarft3_UT_ISW, dv_qr2_dlarft3_isw(1, m, n, a.dev, lda, t.dev, ldt, u.dev))
QR2_T_ENTRY(qr2_dlarft3_ISW_V2T, dv_qr2_dlarft3_isw(0, m, n, a.dev, lda, t.dev, ldt, u.dev))
QR2_T_ENTRY(qr2_dlarft3_ISW_V2T_UT, dv_qr2_dlarft3_isw(1, m, n, a.dev, lda, t.dev, ldt, u.dev))
#undef QR2_T_ENTRY

int qr2_dorgqr3_UT(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau) {
    BEGIN();
    stage_t q = stage_in(Q, ldq, m, n), t = stage_in(T, ldt, n, n), u = stage_in(tau, n, n, 1);
    if (!g_stage_fail) {
        dv_qr2_dorgqr3_UT(m, n, q.dev, ldq, t.dev, ldt, u.dev);
    }
    stage_out(&q, 1); stage_out(&t, 0); stage_out(&u, 0);
    return api_end();
}

/* qr3/src/qr2_dorgqr3_VS2Q.c: the same Q as qr2_dorgqr3_UT */
int qr2_dorgqr3_VS2Q(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau) {
    return qr2_dorgqr3_UT(m, n, Q, ldq, T, ldt, tau);
}

int qr2_aux_dVT2Q(int m, int n, double *Q, int ldq) {
    BEGIN();
    stage_t q = stage_in(Q, ldq, m, n);
    if (!g_stage_fail) {
        dv_qr2_aux_dVT2Q(m, n, q.dev, ldq);
    }
    stage_out(&q, 1);
    return api_end();
}

int qr2_aux_dVS2Q(int m, int n, double *Q, int ldq) {
    BEGIN();
    stage_t q = stage_in(Q, ldq, m, n);
    if (!g_stage_fail) {
        dv_qr2_aux_dVS2Q(m, n, q.dev, ldq);
    }
    stage_out(&q, 1);
    return api_end();
}

int qr2_aux_dmLUinA(int n, double *A, int lda) {
    BEGIN();
    stage_t a = stage_in(A, lda, n, n);
    if (!g_stage_fail) {
        dv_qr2_aux_dmLUinA(n, a.dev, lda);
    }
    stage_out(&a, 1);
    return api_end();
}

#define QR2_LU_ENTRY(NAME, DV)                                      \
    int NAME(int n, double *L, int ldl, double *U, int ldu) {       \
        BEGIN();                                                    \
        const int alias = (L == U);                                 \
        stage_t u = stage_in(U, ldu, n, n);                         \
        stage_t l = alias ? u : stage_in(L, ldl, n, n);             \
        if (!g_stage_fail) {                                         \
            DV(n, alias ? u.dev : l.dev, alias ? ldu : ldl, u.dev, ldu); \
        }                                                            \
        stage_out(&u, 1);                                           \
        if (!alias) stage_out(&l, 0);                               \
        return api_end();                                           \
    }
QR2_LU_ENTRY(qr2_aux_dULTinU, dv_qr2_aux_dULTinU)
QR2_LU_ENTRY(qr2_aux_dUinvLTinU, dv_qr2_aux_dUinvLTinU)
#undef QR2_LU_ENTRY

int qr2_aux_dApUBTinA(int m, int n, double *A, int lda, double *U, int ldu, double *B, int ldb) {
    BEGIN();
    stage_t a = stage_in(A, lda, m, n), u = stage_in(U, ldu, m, m), b = stage_in(B, ldb, n, m);
    if (!g_stage_fail) {
        dv_qr2_aux_dApUBTinA(m, n, a.dev, lda, u.dev, ldu, b.dev, ldb);
    }
    stage_out(&a, 1); stage_out(&u, 0); stage_out(&b, 0);
    return api_end();
}

int qr2_aux_dV2N(int n, double *T, int ldt) {
    BEGIN();
    stage_t t = stage_in(T, ldt, n, n);
    if (!g_stage_fail) {
        dv_qr2_aux_dV2N(n, t.dev, ldt);
    }
    stage_out(&t, 1);
    return api_end();
}

int qr2_aux_dN2T(int n, double *tau, double *T, int ldt) {
    BEGIN();
    stage_t t = stage_in(T, ldt, n, n), u = stage_in(tau, n, n, 1);
    if (!g_stage_fail) {
        dv_qr2_aux_dN2T(n, u.dev, t.dev, ldt);
    }
    stage_out(&t, 1); stage_out(&u, 0);
    return api_end();
}

int qr2_aux_dV2tau(int m, int n, double *A, int lda, double *tau) {
    BEGIN();
    stage_t a = stage_in(A, lda, m, n), u = stage_in(tau, n, n, 1);
    if (!g_stage_fail) {
        dv_qr2_aux_dV2tau(m, n, a.dev, lda, u.dev);
    }
    stage_out(&a, 0); stage_out(&u, 1);
    return api_end();
}

/* LAPACK dlarf through qr3/src/qr2_aux_dlarf_wrapper.c:3-8; work is not
 * used; incv must be >= 1 (-1 otherwise) */
int qr2_aux_dlarf_wrapper(char side, int m, int n, double *V, int incv, double tau, double *C,
                          int ldc, double *work) {
    (void)work;
    if ((side != 'L' && side != 'R') || incv < 1) return -1;
    BEGIN();
    const int len = (side == 'L') ? m : n;
    stage_t v = stage_in(V, incv, 1, len), c = stage_in(C, ldc, m, n);
    if (!g_stage_fail) {
        dv_qr2_aux_dlarf(side, m, n, v.dev, incv, tau, c.dev, ldc);
    }
    stage_out(&v, 0); stage_out(&c, 1);
    return api_end();
}

/* G (n x n, upper triangle written) = A^T A for the m x n A, in the
 * arithmetic of prec (0 or 1 FP64 MFMA, 2 FP32 MFMA, 3 split BF16) */
int lila_dgram(int prec, int m, int n, double *A, int lda, double *G, int ldg) {
    if (prec < 0 || prec > 3 || m < 0 || n < 0) return -1;
    BEGIN();
    stage_t a = stage_in(A, lda, m, n), g = stage_in(G, ldg, n, n);
    if (!g_stage_fail) {
        dv_gram(prec, m, n, a.dev, lda, g.dev, ldg);
    }
    stage_out(&a, 0); stage_out(&g, 1);
    return api_end();
}

/* ---- least squares, generators, status --------------------------------- */

int lila_dgels_qr3(int m, int n, int k, double *A, int lda, double *T, int ldt, double *B, int ldb) {
    BEGIN();
    stage_t a = stage_in(A, lda, m, n), t = stage_in(T, ldt, n, n), b = stage_in(B, ldb, m, k);
    if (!g_stage_fail) {
        dv_ls_qr3(m, n, k, a.dev, lda, t.dev, ldt, b.dev, ldb);
    }
    stage_out(&a, 1); stage_out(&t, 1); stage_out(&b, 1);
    return api_end();
}

int lila_dgemm(char transa, char transb, int m, int n, int k, double alpha, const double *A,
               int lda, const double *B, int ldb, double beta, double *C, int ldc) {
    BEGIN();
    int ta = (transa == 'T' || transa == 't'), tb = (transb == 'T' || transb == 't');
    stage_t a = stage_in((double *)A, lda, ta ? k : m, ta ? m : k);
    stage_t b = stage_in((double *)B, ldb, tb ? n : k, tb ? k : n);
    stage_t c = stage_in(C, ldc, m, n);
    if (!g_stage_fail) {
        ld_gemm_simple(ta, tb, m, n, k, alpha, a.dev, lda, LD_GEN, b.dev, ldb, LD_GEN, beta, c.dev, ldc);
    }
    stage_out(&a, 0); stage_out(&b, 0); stage_out(&c, 1);
    return api_end();
}

/* the same product emulated on int8 MFMA (CRT scheme, kb-bit operands;
 * transb must be 'N'): -1 for a product the emulation does not take */
int lila_dgemm_emu(char transa, char transb, int m, int n, int k, double alpha, const double *A,
                   int lda, const double *B, int ldb, double beta, double *C, int ldc, int kb) {
    int ta = (transa == 'T' || transa == 't'), tb = (transb == 'T' || transb == 't');
    if (tb || m < 0 || n < 0 || k < 0 || k > 65536 || kb < 8 || kb > 53) return -1;
    BEGIN();
    stage_t a = stage_in((double *)A, lda, ta ? k : m, ta ? m : k);
    stage_t b = stage_in((double *)B, ldb, k, n);
    stage_t c = stage_in(C, ldc, m, n);
    int r = 0;
    if (!g_stage_fail) {
        ld_gemm_t g;
        memset(&g, 0, sizeof g);
        g.m = m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
        g.A = a.dev; g.lda = lda; g.ta = ta; g.sa = LD_GEN;
        g.B = b.dev; g.ldb = ldb; g.tb = 0; g.sb = LD_GEN;
        g.C = c.dev; g.ldc = ldc;
        if (k == 0 || alpha == 0.0) ld_gemm(&g);
        else r = ld_gemm_emu(&g, kb);
    }
    stage_out(&a, 0); stage_out(&b, 0); stage_out(&c, r == 0);
    const int e = api_end();
    return e != 0 ? e : r;
}

int lila_fill_uniform(unsigned long long seed, int m, int n, double *A, int lda, double lo, double hi) {
    BEGIN();
    stage_t a = stage_in(A, lda, m, n);
    if (!g_stage_fail) {
        ld_fill_uniform(seed, m, n, a.dev, lda, lo, hi);
    }
    stage_out(&a, 1);
    return api_end();
}

int lila_fill_normal(unsigned long long seed, int m, int n, double *A, int lda) {
    BEGIN();
    stage_t a = stage_in(A, lda, m, n);
    if (!g_stage_fail) {
        ld_fill_normal(seed, m, n, a.dev, lda);
    }
    stage_out(&a, 1);
    return api_end();
}

int lila_get_info(void) {
    if (ld_init() != 0) return -1;
    ld_sync();
    int r = ld_get_info();
    ld_clear_info();
    return r;
}

int lila_sync(void) {
    if (ld_init() != 0) return -1;
    ld_sync();
    return ld_get_info();
}

/* ---- batched least squares -------------------------------------------- */

lila_batch_plan *lila_dgels_batched_plan(int nprob, const int *m, const int *n, const int *nrhs,
                                         double *const *A, const int *lda, double *const *B,
                                         const int *ldb) {
    if (ld_init() != 0) {
        fprintf(stderr, "lila_hip: no usable HIP device; this library has no CPU path\n");
        return NULL;
    }
    for (int p = 0; p < nprob; p++) {
        if (!ld_is_device_ptr(A[p]) || (nrhs[p] > 0 && !ld_is_device_ptr(B[p]))) {
            int e = 0, t = 0;
            ld_ptr_diag(ld_is_device_ptr(A[p]) ? (const void *)B[p] : (const void *)A[p], &e, &t);
            fprintf(stderr,
                    "lila_hip: batched plan needs device pointers (problem %d: %p / %p, "
                    "hipPointerGetAttributes status %d, memory type %d)\n",
                    p, (void *)A[p], (void *)B[p], e, t);
            return NULL;
        }
    }
    return ld_batch_plan_create(nprob, m, n, nrhs, A, lda, B, ldb);
}

int lila_dgels_batched_run(lila_batch_plan *plan, int use_graph) {
    if (!plan) return -1;
    return ld_batch_plan_run(plan, use_graph);
}

double lila_dgels_batched_flops(const lila_batch_plan *plan) { return ld_batch_plan_flops(plan); }

void lila_dgels_batched_destroy(lila_batch_plan *plan) { ld_batch_plan_destroy(plan); }

int lila_dgels_batched(int nprob, const int *m, const int *n, const int *nrhs, double *const *A,
                       const int *lda, double *const *B, const int *ldb) {
    BEGIN();
    if (nprob == 0) return 0;
    if (nprob < 0) return -1;
    int ndev = 0;
    for (int p = 0; p < nprob; p++) ndev += ld_is_device_ptr(A[p]);
    if (ndev != 0 && ndev != nprob) {
        fprintf(stderr, "lila_hip: lila_dgels_batched mixes host and device pointers\n");
        return -1;
    }
    stage_t *sa = (stage_t *)calloc((size_t)nprob, sizeof(stage_t));
    stage_t *sb = (stage_t *)calloc((size_t)nprob, sizeof(stage_t));
    double **da = (double **)malloc(sizeof(double *) * (size_t)nprob);
    double **db = (double **)malloc(sizeof(double *) * (size_t)nprob);
    int r = -1;
    if (!sa || !sb || !da || !db) goto done;
    /* stage every problem; a failed staging allocation launches nothing */
    for (int p = 0; p < nprob && !g_stage_fail; p++) {
        sa[p] = stage_in(A[p], lda[p], m[p], n[p]);
        sb[p] = stage_in(B[p], ldb[p], m[p], nrhs[p]);
        da[p] = sa[p].dev;
        db[p] = sb[p].dev;
    }
    if (!g_stage_fail) {
        ld_batch_plan *pl = ld_batch_plan_create(nprob, m, n, nrhs, da, lda, db, ldb);
        if (pl) {
            r = ld_batch_plan_run(pl, 0);
            ld_batch_plan_destroy(pl);  /* waits for the launch */
        }
    }
    for (int p = 0; p < nprob; p++) {
        stage_out(&sa[p], r == 0);
        stage_out(&sb[p], r == 0);
    }
    if (g_stage_fail) r = api_end();
    else if (r == 0) r = api_end();
done:
    free(sa); free(sb); free(da); free(db);
    return r;
}

/* no allocation: the problem table, queue order and counter in the
 * library's static device scratch (batched.hip); device pointers only;
 * asynchronous on the library stream */
int lila_dgels_batched_noalloc(int nprob, const int *m, const int *n, const int *nrhs,
                               double *const *A, const int *lda, double *const *B, const int *ldb) {
    if (nprob == 0) return 0;
    if (nprob < 0 || !m || !n || !nrhs || !A || !lda || !B || !ldb) return -1;
    if ((size_t)nprob > ld_batch_noalloc_max()) {
        fprintf(stderr, "lila_hip: lila_dgels_batched_noalloc takes at most %zu problems (got %d)\n",
                ld_batch_noalloc_max(), nprob);
        return -1;
    }
    for (int p = 0; p < nprob; p++) {
        if (!ld_is_device_ptr(A[p]) || (nrhs[p] > 0 && !ld_is_device_ptr(B[p]))) {
            fprintf(stderr, "lila_hip: lila_dgels_batched_noalloc needs device pointers (problem %d)\n",
                    p);
            return -1;
        }
    }
    return ld_batch_noalloc_run(nprob, m, n, nrhs, A, lda, B, ldb);
}

void lila_set_stream(void *s) { ld_set_stream(s); }
void *lila_get_stream(void) { return ld_stream(); }

/* ---- workspace queries (closed forms of lila-c/lila_query_*.c) ---------- */

int lila_query_dormqrf_w03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T,
                           int ldt, double *work, int lwork) {
    (void)j; (void)A; (void)lda; (void)T; (void)ldt; (void)work; (void)lwork;
    return dq_dormqrf_w03(m, n, k, i, mt);
}

int lila_query_dormqrbz_w03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *Q,
                            int ldq, double *T, int ldt, double *work, int lwork) {
    (void)j; (void)A; (void)lda; (void)Q; (void)ldq; (void)T; (void)ldt; (void)work; (void)lwork;
    return dq_dormqrbz_w03(m, n, k, i, mt);
}

int lila_query_dgeqrf_w03_mt(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                             double *T, int ldt, double *Q, int ldq, double *work, int lwork) {
    (void)A; (void)lda; (void)T; (void)ldt; (void)Q; (void)ldq; (void)work;
    return dq_dgeqrf_w03_mt(lila_param, m, n, i, mt, lwork);
}

int lila_query_dgeqrf_w03_mt_hh(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                double *T, int ldt, double *Q, int ldq, double *work, int lwork) {
    (void)A; (void)lda; (void)T; (void)ldt; (void)Q; (void)ldq; (void)work;
    int p[4] = {lila_param[0], 1, lila_param[2], lila_param[3]};
    return dq_dgeqrf_w03_mt(p, m, n, i, mt, lwork);
}

int lila_query_dgeqrf_w03_mt_hr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                                double *Q, int ldq, double *work, int lwork) {
    (void)A; (void)lda; (void)T; (void)ldt; (void)Q; (void)ldq; (void)work;
    int p[4] = {0, 2, 0, 0};
    return dq_dgeqrf_w03_mt(p, m, n, i, mt, lwork);
}

int lila_query_dgeqrf_w03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                               double *Q, int ldq, double *work, int lwork) {
    (void)m; (void)i; (void)mt; (void)A; (void)lda; (void)T; (void)ldt; (void)Q; (void)ldq;
    (void)work; (void)lwork;
    return n * n;
}

int lila_query_dgeqr2_w03_hr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                             double *Q, int ldq, double *work, int lwork) {
    (void)m; (void)i; (void)mt; (void)A; (void)lda; (void)T; (void)ldt; (void)Q; (void)ldq;
    (void)work; (void)lwork;
    return n;
}

int lila_query_dgeqr2_w03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                            double *Q, int ldq, double *work, int lwork) {
    (void)m; (void)i; (void)mt; (void)A; (void)lda; (void)T; (void)ldt; (void)Q; (void)ldq;
    (void)work; (void)lwork;
    return n * n;
}

int lila_query_dgeqr2_w03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                            double *Q, int ldq, double *work, int lwork) {
    (void)m; (void)i; (void)mt; (void)A; (void)lda; (void)T; (void)ldt; (void)Q; (void)ldq;
    (void)work; (void)lwork;
    return n * n;
}

int lila_query_ormhr2_w03_hr(int m, int n, int i, int j, int l, int mt, double *A, int lda, double *T,
                             int ldt, double *Q, int ldq, double *work, int lwork, double *S) {
    (void)m; (void)l; (void)mt; (void)A; (void)lda; (void)T; (void)ldt; (void)Q; (void)ldq;
    (void)work; (void)lwork; (void)S;
    return n * (j - i) + n;
}

int lila_query_dorgh2_3(int m, int n, double *A, int lda, double *T, int ldt, double *Q, int ldq,
                        double *work, int lwork, double *S) {
    (void)m; (void)n; (void)A; (void)lda; (void)T; (void)ldt; (void)Q; (void)ldq; (void)work;
    (void)lwork; (void)S;
    return 0;
}

int lila_query_dgeqrf_w03_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                    double *T, int ldt, double *Q, int ldq, double *work, int lwork) {
    (void)A; (void)lda; (void)T; (void)ldt; (void)Q; (void)ldq; (void)work;
    return dq_dgeqrf_w03_recursive(lila_param, m, n, i, mt, lwork);
}

int lila_query_dgeqrf_w03_levelx(int *lila_param, int n_lvl, int i_lvl, int *nb_lvl, int m, int n,
                                 int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                                 int ldq, double *work, int lwork) {
    (void)A; (void)lda; (void)T; (void)ldt; (void)Q; (void)ldq; (void)work; (void)lwork;
    return dq_dgeqrf_w03_levelx(lila_param, n_lvl, i_lvl, nb_lvl, m, n, i, mt);
}

/* lila-c/lila_query_dgeqrf_ker_recursive.c: the recursive results are
 * overwritten before use there, so only the leaf (n+i, n*n) or split
 * (nb1*nb2) terms of the top call count */
int lila_query_dgeqrf_ker_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                    double *T, int ldt, double *Q, int ldq, double *work, int lwork) {
    (void)m; (void)mt; (void)A; (void)lda; (void)T; (void)ldt; (void)Q; (void)ldq; (void)work;
    int w = lwork;
    if (n <= lila_param[3]) {
        if (w < n + i) w = n + i;
        if (w < n * n) w = n * n;
    } else {
        int nb1 = n / 2, nb2 = n - nb1;
        if (w < nb1 * nb2) w = nb1 * nb2;
    }
    return w;
}

/* lila-c/lila_query_dgeqrf_LAPACK_appendcols.c */
int lila_query_dgeqrf_LAPACK_appendcols(int m, int k, int n, int mt, double *A, int lda, double *Q,
                                        int ldq, double *T, int ldt, double *work, int lwork) {
    (void)m; (void)mt; (void)A; (void)lda; (void)Q; (void)ldq; (void)T; (void)ldt; (void)work;
    int w = lwork;
    if (w < n * k) w = n * k;
    if (w < n * n) w = n * n;
    return w;
}
