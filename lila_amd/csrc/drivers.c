/*
 * drivers.c -- host-side orchestration of the LILA hot path on one MI355X.
 *
 * Plain C.  Each dv_* routine keeps the call tree and the global-index /
 * mt-blocked-T pointer arithmetic of the reference routine it replaces
 * (cited per function), and issues device work through lila_dev.h: the
 * BLAS-3 steps become FP64 MFMA GEMM launches that read the Householder V
 * (unit lower) and T (upper) in place through operand structure masks, the
 * panel becomes one fused launch, and the Cholesky-QR / reconstruction
 * steps become TRSM / SYRK / Cholesky / sign-LU launches.  All pointers are
 * device pointers; all work goes onto the library stream.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lila_dev.h"
#include "lila_drivers.h"

#define N_ 0
#define T_ 1

static double *wsd(size_t n) { return (double *)ld_ws_alloc(n * sizeof(double)); }

static void gemm(int ta, int tb, int m, int n, int k, double alpha, const double *A, int lda,
                 int sa, const double *B, int ldb, int sb, double beta, double *C, int ldc) {
    ld_gemm_simple(ta, tb, m, n, k, alpha, A, lda, sa, B, ldb, sb, beta, C, ldc);
}

/* ---- block reflector pieces ------------------------------------------- */

/* W = op(T) (op(A) op(B)) for a small T (op(T) = T^T when transT): one
 * split product plus one launch that sums it and multiplies by op(T) when
 * the product is split over K, else the product and a triangular GEMM */
static void gemm_then_tri(int ta, int tb, int m, int n, int k, const double *A, int lda, int sa,
                          const double *B, int ldb, int sb, const double *T, int ldt, int transT,
                          double *W, int ldw) {
    ld_gemm_t g;
    memset(&g, 0, sizeof g);
    g.m = m; g.n = n; g.k = k; g.alpha = 1.0; g.beta = 0.0;
    g.A = A; g.lda = lda; g.ta = ta; g.sa = sa;
    g.B = B; g.ldb = ldb; g.tb = tb; g.sb = sb;
    if (ld_gemm_tmul(&g, T, ldt, transT, W, ldw) == 0) return;
    size_t mark = ld_ws_mark();
    double *G = wsd((size_t)m * n);
    gemm(ta, tb, m, n, k, 1.0, A, lda, sa, B, ldb, sb, 0.0, G, m);
    gemm(transT ? T_ : N_, N_, m, n, m, 1.0, T, ldt, LD_UPPER, G, m, LD_GEN, 0.0, W, ldw);
    ld_ws_release(mark);
}

/* C (rows [0, ml) of the block) <- (I - V T V^T)^T C for one block of vb
 * reflectors: W = V^T C; W = T^T W; C -= V W. */
static void apply_block_qt(int ml, int n, int vb, const double *V, int ldv, const double *Tb,
                           int ldt, double *C, int ldc) {
    if (vb <= 0 || n <= 0) return;
    size_t mark = ld_ws_mark();
    double *W2 = wsd((size_t)vb * n);
    gemm_then_tri(T_, N_, vb, n, ml, V, ldv, LD_VUNIT, C, ldc, LD_GEN, Tb, ldt, 1, W2, vb);
    gemm(N_, N_, ml, n, vb, -1.0, V, ldv, LD_VUNIT, W2, vb, LD_GEN, 1.0, C, ldc);
    ld_ws_release(mark);
}

/* C <- H^T C for the k reflectors starting at global i, mt-block by
 * mt-block; C row 0 is global row i.  Body shared by
 * lila-c/lila_dormqrf_w03.c:3-81 and lila-c/lila_dormqrf_z03.c:3-72: block b
 * starts at global b0, has vb = min(mt - b0%mt, remaining) reflectors and
 * its T at T + b0%mt + b0*ldt (the walk of lila-c/lila_dormqrf_w03.c:63,71). */
static void apply_qt(int m, int n, int k, int i, int mt, const double *A, int lda,
                     const double *T, int ldt, double *C, int ldc) {
    if (k <= 0 || n <= 0) return;
    int b0 = i, done = 0, vb = mt - i % mt;
    if (vb > k) vb = k;
    while (done < k) {
        apply_block_qt(m - b0, n, vb, A + b0 + (size_t)b0 * lda, lda, T + b0 % mt + (size_t)b0 * ldt,
                       ldt, C + (b0 - i), ldc);
        done += vb;
        b0 += vb;
        vb = (k - done >= mt) ? mt : (k - done);
    }
}

int dv_dormqrf_w03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T, int ldt) {
    apply_qt(m, n, k, i, mt, A, lda, T, ldt, A + i + (size_t)j * lda, lda);
    return 0;
}

int dv_dormqrf_z03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T, int ldt,
                   double *B, int ldb) {
    apply_qt(m, n, k, i, mt, A, lda, T, ldt, B + i + (size_t)j * ldb, ldb);
    return 0;
}

/* Q(:, j:j+n) <- H_i ... H_{i+k-1} [0; Q_new], mt-blocks walked backwards;
 * lila-c/lila_dormqrbz_w03.c:3-66.  Per block: W = V2^T Q2; W = T W;
 * Q_block = [0; Q2] - V W (one GEMM; the top vb rows take beta = 0, which is
 * the dlacpy + trmm of :27-28). */
int dv_dormqrbz_w03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *Q, int ldq,
                    double *T, int ldt) {
    if (k <= 0 || n <= 0) return 0;
    int cap = k / (mt > 0 ? mt : 1) + 2;
    int *start = (int *)malloc((size_t)cap * sizeof(int));
    int nblk = 0, b0 = i, vb = mt - i % mt;
    if (vb > k) vb = k;
    while (b0 < i + k) {
        start[nblk++] = b0;
        b0 += vb;
        vb = (i + k - b0 >= mt) ? mt : (i + k - b0);
    }
    for (int b = nblk - 1; b >= 0; b--) {
        int s = start[b], e = (b + 1 < nblk) ? start[b + 1] : i + k;
        int w = e - s, ml = m - s;
        const double *Vb = A + s + (size_t)s * lda;
        const double *Tb = T + s % mt + (size_t)s * ldt;
        double *Qb = Q + s + (size_t)j * ldq;
        size_t mark = ld_ws_mark();
        double *W2 = wsd((size_t)w * n);
        gemm_then_tri(T_, N_, w, n, ml - w, Vb + w, lda, LD_GEN, Qb + w, ldq, LD_GEN, Tb, ldt, 0, W2,
                      w);
        ld_gemm_t g;
        memset(&g, 0, sizeof g);
        g.m = ml; g.n = n; g.k = w; g.alpha = -1.0; g.beta = 1.0;
        g.A = Vb; g.lda = lda; g.ta = 0; g.sa = LD_VUNIT;
        g.B = W2; g.ldb = w; g.tb = 0; g.sb = LD_GEN;
        g.C = Qb; g.ldc = ldq; g.zero_rows = w;
        ld_gemm(&g);
        ld_ws_release(mark);
    }
    free(start);
    return 0;
}

/* Join T of the reflectors [i, i+vb) with the wb earlier ones of the same
 * mt-block: T_ji = -T_jj (V_j^T V_i) T_ii; lila-c/lila_dlarft_connect_w03.c:3-32
 * (the wb/vb rule of :8-11; the copy + trmm + gemm of :19-26 is one GEMM with
 * V_i read as unit lower). */
int dv_dlarft_connect_w03(int m, int n, int i, int j, int mt, double *A, int lda, double *T, int ldt) {
    int vb = mt - i % mt;
    if (vb > n) vb = n;
    int wb = ((j / mt) == (i / mt)) ? (i - j) : (i % mt);
    if (wb == mt) wb = 1;
    if (wb <= 0 || vb <= 0) return 0;
    const double *Aii = A + i + (size_t)i * lda;
    const double *Aij = A + i + (size_t)(i - wb) * lda;
    const double *Tki = T + i % mt + (size_t)i * ldt;
    double *Tji = T + (i % mt - wb) + (size_t)i * ldt;
    const double *Tjj = T + (i % mt - wb) + (size_t)(i - wb) * ldt;
    size_t mark = ld_ws_mark();
    double *H = wsd((size_t)wb * vb);
    gemm_then_tri(T_, N_, wb, vb, m - i, Aij, lda, LD_GEN, Aii, lda, LD_VUNIT, Tjj, ldt, 0, H, wb);
    gemm(N_, N_, wb, vb, vb, -1.0, H, wb, LD_GEN, Tki, ldt, LD_UPPER, 0.0, Tji, ldt);
    ld_ws_release(mark);
    return 0;
}

/* ---- panels ------------------------------------------------------------- */

static void reconstruct_block(int ml, int n, double *A, int lda, double *Tq, int ldt, double *Q,
                              int ldq, int mq, double *S);

/* TSQR of a tall ml x nb panel (nb <= 32), the tall-panel algorithm of
 * tsqr/dlatsqr.f (slab QRs whose R factors are factored again) arranged as
 * a tree: every level's slabs are independent QRs (one workgroup each), the
 * stacked R's form the next level, until one slab is left.  The explicit Q
 * is built back down the tree (tsqr/dlamtsqr.f's application to [I; 0]),
 * and Q with the final R is turned into dgeqr3's V, R and T by the
 * sign-chosen LU reconstruction (lila-c/lila_dorgh2_3.c:8-51).  Householder
 * arithmetic throughout: stable for any conditioning.  Q, when given,
 * receives the explicit panel Q. */
#define TSQR_MAXLEV 16
static void tsqr_panel(int ml, int nb, double *A, int lda, double *T, int ldt, double *Q, int ldq) {
    size_t mark = ld_ws_mark();
    double *X[TSQR_MAXLEV + 1], *Ts[TSQR_MAXLEV];
    int M[TSQR_MAXLEV + 1], ldx[TSQR_MAXLEV + 1], S[TSQR_MAXLEV];
    int L = 0;
    /* the tree works on a copy: A stays untouched for the column loop below */
    X[0] = wsd((size_t)ml * nb);
    ldx[0] = ml;
    M[0] = ml;
    ld_lacpy('A', ml, nb, A, lda, X[0], ml);
    for (;;) {
        S[L] = ld_tsqr_slabs(M[L], nb);
        Ts[L] = wsd((size_t)S[L] * 32 * 32);
        M[L + 1] = S[L] * nb;
        ldx[L + 1] = M[L + 1];
        X[L + 1] = wsd((size_t)M[L + 1] * nb);
        ld_tsqr_leaf(M[L], nb, X[L], ldx[L], S[L], Ts[L], X[L + 1], ldx[L + 1]);
        L++;
        if (S[L - 1] == 1 || L == TSQR_MAXLEV) break;
    }
    /* X[L] (nb x nb) holds the final R; the explicit Q down the levels */
    double *Qn = NULL;
    int ldqn = 0;
    for (int l = L - 1; l >= 0; l--) {
        double *Qo;
        int ldo;
        if (l == 0) {
            Qo = Q ? Q : wsd((size_t)ml * nb);
            ldo = Q ? ldq : ml;
        } else {
            Qo = wsd((size_t)M[l] * nb);
            ldo = M[l];
        }
        ld_tsqr_expand(M[l], nb, S[l], X[l], ldx[l], Ts[l], Qn, ldqn, Qo, ldo);
        Qn = Qo;
        ldqn = ldo;
    }
    /* A reflector with tau = 0 (x = 0: dgeqr3 leaves the column alone) or
     * within 1e-10 of 0 or 2 shows as |Q_kk| within 1e-10 of 1, which the
     * reconstruction cannot reproduce (it always reflects): such panels take
     * dgeqr3's column-by-column loop on the untouched A, as the fused
     * kernels' acceptance test does */
    {
        double *dq = wsd((size_t)nb), hq[32];
        ld_T2tau(nb, 0, nb + 1, Qn, ldqn, dq);
        ld_memcpy_d2h(hq, dq, sizeof(double) * (size_t)nb);
        int degenerate = 0;
        for (int k = 0; k < nb; k++) degenerate |= !(fabs(hq[k]) <= 1.0 - 1e-10);
        if (degenerate) {
            ld_ws_release(mark);
            ld_panel_qr_exact(ml, nb, A, lda, T, ldt, Q, ldq);
            return;
        }
    }
    /* reconstruction inputs: R on top of A, Q2 below, Q1 in T */
    ld_lacpy('U', nb, nb, X[L], ldx[L], A, lda);
    if (ml > nb) ld_lacpy('A', ml - nb, nb, Qn + nb, ldqn, A + nb, lda);
    ld_lacpy('A', nb, nb, Qn, ldqn, T, ldt);
    double *Sd = wsd((size_t)nb);
    reconstruct_block(ml, nb, A, lda, T, ldt, Q, ldq, Q ? ml : 0, Sd);
    ld_ws_release(mark);
}

/* one panel of at most 32 columns: the fused kernel, or for panels taller
 * than its LDS slabs TSQR (column-by-column mode), or the streamed
 * Cholesky-QR2 kernel with TSQR for the panels it declines (ill-conditioned:
 * a Cholesky or LU acceptance test failed) */
static void panel(int ml, int nb, double *A, int lda, double *T, int ldt, double *Q, int ldq) {
    if (!ld_panel_is_tall(ml, nb)) {
        ld_panel_qr(ml, nb, A, lda, T, ldt, Q, ldq);
        return;
    }
    if (ld_tune()->panel_mode == 1 && ld_panel_qr_tall_cqr(ml, nb, A, lda, T, ldt, Q, ldq) == 0) return;
    tsqr_panel(ml, nb, A, lda, T, ldt, Q, ldq);
}

/* Recursive panel QR (Elmroth-Gustavson), lila-c/dgeqr3.c:3-114: the left
 * half recursively, its reflectors applied to the right half (:44-70), the
 * right half recursively (:75), T12 = -T11 (V1^T V2) T22 (:95-108).  Widths
 * up to 32 go to the fused panel kernel. */
/* the panel width: widths up to it go to one panel launch (the autotuner's
 * choice, 8..32) */
static int node_kernels(void) { return ld_tune()->node_fused; }

static int panel_width(void) {
    const int w = ld_tune()->panel_nb;
    return (w >= 8 && w <= LD_PANEL_MAX) ? w : LD_PANEL_MAX;
}

int dv_dgeqr3(int m, int n, double *A, int lda, double *T, int ldt) {
    if (n <= 0) return 0;
    if (ld_tune()->panel_nb == 64 && n > LD_PANEL_MAX && n <= 64 &&
        ld_panel_qr64(m, n, A, lda, T, ldt) == 0)
        return 0;
    if (n <= panel_width()) {
        panel(m, n, A, lda, T, ldt, NULL, 0);
        return 0;
    }
    int n1 = n / 2, n2 = n - n1;
    const int node = n1 <= 64 && n2 <= 64 && node_kernels();
    dv_dgeqr3(m, n1, A, lda, T, ldt);
    if (!node || ld_node_update64(m, n1, n2, A, lda, T, ldt, A + (size_t)n1 * lda, lda) != 0)
        apply_block_qt(m, n2, n1, A, lda, T, ldt, A + (size_t)n1 * lda, lda);
    dv_dgeqr3(m - n1, n2, A + n1 + (size_t)n1 * lda, lda, T + n1 + (size_t)n1 * ldt, ldt);
    if (node && ld_node_join64(m - n1, n1, n2, A + n1, lda, A + n1 + (size_t)n1 * lda, lda, T, ldt,
                               T + n1 + (size_t)n1 * ldt, ldt, T + (size_t)n1 * ldt, ldt) == 0)
        return 0;
    size_t mark = ld_ws_mark();
    double *H = wsd((size_t)n1 * n2);
    gemm_then_tri(T_, N_, n1, n2, m - n1, A + n1, lda, LD_GEN, A + n1 + (size_t)n1 * lda, lda,
                  LD_VUNIT, T, ldt, 0, H, n1);
    gemm(N_, N_, n1, n2, n2, -1.0, H, n1, LD_GEN, T + n1 + (size_t)n1 * ldt, ldt, LD_UPPER, 0.0,
         T + (size_t)n1 * ldt, ldt);
    ld_ws_release(mark);
    return 0;
}

/* explicit panel Q = [I;0] - V (T V1^T) (lila-c/lila_dgeqr2_w03_3.c:19-30) */
static void panel_q(int ml, int n, const double *V, int ldv, const double *T, int ldt, double *Q,
                    int ldq) {
    size_t mark = ld_ws_mark();
    double *X = wsd((size_t)n * n);
    gemm(N_, T_, n, n, n, 1.0, T, ldt, LD_UPPER, V, ldv, LD_VUNIT, 0.0, X, n);
    gemm(N_, N_, ml, n, n, -1.0, V, ldv, LD_VUNIT, X, n, LD_UPPER, 0.0, Q, ldq);
    ld_add_diag(n, Q, ldq, 1.0);
    ld_ws_release(mark);
}

/* dgeqr3 + explicit panel Q; lila-c/lila_dgeqr2_w03_3.c:3-34 */
int dv_dgeqr2_w03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                    int ldq) {
    int ml = m - i;
    double *Aii = A + i + (size_t)i * lda, *Qii = Q + i + (size_t)i * ldq;
    double *Tki = T + i % mt + (size_t)i * ldt;
    if (n <= panel_width()) {
        panel(ml, n, Aii, lda, Tki, ldt, Qii, ldq);
    } else {
        dv_dgeqr3(ml, n, Aii, lda, Tki, ldt);
        panel_q(ml, n, Aii, lda, Tki, ldt, Qii, ldq);
    }
    return 0;
}

/* LAPACK-style panel (dgeqrf + dlarft + dorgqr; lila-c/lila_dgeqr2_w03_l.c).
 * On the GPU the Householder vectors, tau and T of dgeqrf + dlarft are
 * exactly what the recursive panel produces, so it shares that path. */
int dv_dgeqr2_w03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                    int ldq) {
    return dv_dgeqr2_w03_3(m, n, i, mt, A, lda, T, ldt, Q, ldq);
}

/* Sign-chosen LU reconstruction of one block (lila-c/lila_dorgh2_3.c:8-51):
 * A (ml rows) holds R on top and Q2 below, Tq the n x n Q1 copy, Q the
 * explicit Q (mq rows) whose columns are flipped; S (device) gets +-1. */
static void reconstruct_block(int ml, int n, double *A, int lda, double *Tq, int ldt, double *Q,
                              int ldq, int mq, double *S) {
    ld_lu_sign(n, Tq, ldt, S);
    if (ml > n) {
        ld_flip_cols(ml - n, n, A + n, lda, S);
        ld_trsm('R', 'U', 'N', 'N', ml - n, n, 1.0, Tq, ldt, A + n, lda);
    }
    ld_flip_rows_upper(n, A, lda, S);
    ld_flip_cols(mq, n, Q, ldq, S);
    ld_move_lower(n, Tq, ldt, A, lda);
    ld_trsm('R', 'L', 'T', 'U', n, n, -1.0, A, lda, Tq, ldt);
}

int dv_dorgh2_3(int m, int n, double *A, int lda, double *T, int ldt, double *Q, int ldq, double *S) {
    reconstruct_block(m, n, A, lda, T, ldt, Q, ldq, m, S);
    return 0;
}

/* global-index form, mt layout (storage/c2/lila_dorgh2_w03.c:3-65) */
int dv_dorgh2(int m, int n, int i, int j, int mt, double *A, int lda, double *T, int ldt, double *Q,
              int ldq, double *S) {
    reconstruct_block(m - j, n, A + j + (size_t)j * lda, lda, T + j % mt + (size_t)j * ldt, ldt,
                      Q + i + (size_t)j * ldq, ldq, m - i, S + j);
    return 0;
}

/* X = L^-1 B for the unit lower triangle L of the reconstructed columns
 * [i, i + jb) (rows i.., ld lda), B jb x nrhs (ld jb), by the blocks the
 * reconstruction made (the first mt - i % mt columns, then mt each): per
 * block X_k = Linv_k B_k with the block's cached inverse (Linv: mt x jb,
 * block k's inverse in its columns), then the rows below -= L_below X_k.
 * B is overwritten by the partial sums; X receives the solution. */
static void unit_lower_solve_blocked(int jb, int nrhs, int i, int mt, const double *A, int lda,
                                     const double *Linv, double *B, double *X) {
    for (int c = 0, w = mt - i % mt; c < jb; c += w, w = mt) {
        if (w > jb - c) w = jb - c;
        gemm(N_, N_, w, nrhs, w, 1.0, Linv + (size_t)c * mt, mt, LD_GEN, B + c, jb, LD_GEN, 0.0, X + c,
             jb);
        if (c + w < jb)
            gemm(N_, N_, jb - c - w, nrhs, w, -1.0, A + i + c + w + (size_t)(i + c) * lda, lda, LD_GEN,
                 X + c, jb, LD_GEN, 1.0, B + c + w, jb);
    }
}

/* blocked reconstruction step, lila-c/lila_ormhr2_w03_hr.c:3-84 (S kept
 * apart from the workspace); Linv (or NULL: a plain TRSM, :33) holds the
 * inverses of the earlier blocks' unit lower diagonal blocks */
static int ormhr2_step(int m, int n, int i, int j, int mt, double *A, int lda, double *T, int ldt,
                       double *Q, int ldq, double *S, const double *Linv) {
    int jb = j - i;
    double *Tjj = T + j % mt + (size_t)j * ldt;
    double *Ajj = A + j + (size_t)j * lda;
    size_t mark = ld_ws_mark();
    if (jb > 0) ld_flip_rows(jb, n, A + i + (size_t)j * lda, lda, S + i);
    double *W2 = wsd((size_t)(jb > 0 ? jb : 1) * n);
    if (jb > 0) ld_lacpy('A', jb, n, Q + i + (size_t)j * ldq, ldq, W2, jb);
    ld_lacpy('A', n, n, Q + j + (size_t)j * ldq, ldq, Tjj, ldt);
    ld_lacpy('A', m - j - n, n, Q + j + n + (size_t)j * ldq, ldq, Ajj + n, lda);
    if (jb > 0) {
        if (Linv) {
            double *X = wsd((size_t)jb * n);
            unit_lower_solve_blocked(jb, n, i, mt, A, lda, Linv, W2, X);
            W2 = X;
        } else {
            ld_trsm('L', 'L', 'N', 'U', jb, n, 1.0, A + i + (size_t)i * lda, lda, W2, jb);
        }
        gemm(N_, N_, n, n, jb, -1.0, A + j + (size_t)i * lda, lda, LD_GEN, W2, jb, LD_GEN, 1.0, Tjj, ldt);
        gemm(N_, N_, m - j - n, n, jb, -1.0, A + j + n + (size_t)i * lda, lda, LD_GEN, W2, jb, LD_GEN,
             1.0, Ajj + n, lda);
    }
    reconstruct_block(m - j, n, Ajj, lda, Tjj, ldt, Q + i + (size_t)j * ldq, ldq, m - i, S + j);
    ld_ws_release(mark);
    return 0;
}

/* lila-c/lila_ormhr2_w03_hr.c:3-84 (the API entry: no cached inverses) */
int dv_ormhr2_w03_hr(int m, int n, int i, int j, int l, int mt, double *A, int lda, double *T, int ldt,
                     double *Q, int ldq, double *S) {
    (void)l;
    return ormhr2_step(m, n, i, j, mt, A, lda, T, ldt, Q, ldq, S, NULL);
}

/* blocked reconstruction driver (loop of lila-c/lila_dgeqrf_w03_mt_hr.c:23-29);
 * Sd: device doubles indexed by global column.  Each step solves with the
 * unit lower triangle of all earlier blocks' V; their diagonal blocks do not
 * change after their own step, so each is inverted once (a w x w TRSM on
 * the identity) and the later solves run by blocks of mt rows, two GEMMs
 * per block, instead of a TRSM recursion down to 64-row leaves per step. */
int dv_dorghr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q, int ldq,
              double *Sd) {
    size_t mark = ld_ws_mark();
    double *Linv = n > mt - i % mt ? wsd((size_t)mt * n) : NULL;
    int j = i, vb = mt - i % mt;
    if (vb > n) vb = n;
    while (vb > 0) {
        ormhr2_step(m, vb, i, j, mt, A, lda, T, ldt, Q, ldq, Sd, Linv);
        const int next = (j + vb + mt >= i + n) ? (i + n - j - vb) : mt;
        if (Linv && next > 0) {
            double *Lk = Linv + (size_t)(j - i) * mt;
            ld_laset('A', vb, vb, 0.0, 1.0, Lk, mt);
            ld_trsm('L', 'L', 'N', 'U', vb, vb, 1.0, A + j + (size_t)j * lda, lda, Lk, mt);
        }
        j += vb;
        vb = next;
    }
    ld_ws_release(mark);
    return 0;
}

/* one Cholesky-QR pass: Q = A R^{-1}, R = chol(A^T A) in the top of A */
/* G (upper) = X^T X in the arithmetic of `prec`: FP64 MFMA (0, 1; the
 * tall-Gram kernel for n <= 256, the SYRK form of the GEMM above), FP32
 * MFMA with operands rounded to FP32 (2), 3-term split-BF16 MFMA with FP32
 * accumulation (3, n <= 256; FP32 MFMA above) */
int dv_gram(int prec, int ml, int n, const double *X, int ldx, double *G, int ldg) {
    if (prec == 3 && n <= 256) return ld_gram_bf16x3(ml, n, X, ldx, G, ldg);
    if (prec <= 1 && n <= 256) return ld_gram_f64(ml, n, X, ldx, G, ldg);
    ld_tune_t *tu = ld_tune();
    const int old = tu->gemm_lowp;
    if (prec >= 2) tu->gemm_lowp = 1;
    ld_syrk_ut(n, ml, 1.0, X, ldx, 0.0, G, ldg);
    tu->gemm_lowp = old;
    return 0;
}

/* Cholesky-QR of the hr paths (lila-c/lila_dgeqrf_w03_mt_hr.c:18-21: dsyrk,
 * dpotrf, dtrsm): Q = A R^-1 with R on and above A's diagonal.  With the
 * Gram precision set to 1, 2 or 3 it runs twice (Cholesky-QR2), the first
 * Gram in FP64, FP32 MFMA or split BF16 and the second in FP64:
 * Q = (A R1^-1) R2^-1, R = R2 R1. */
static void cholqr(int ml, int n, double *Aii, int lda, double *Qii, int ldq) {
    const int prec = ld_tune()->gram_prec;
    ld_lacpy('A', ml, n, Aii, lda, Qii, ldq);
    dv_gram(prec, ml, n, Qii, ldq, Aii, lda);
    ld_potrf_u(n, Aii, lda);
    ld_trsm('R', 'U', 'N', 'N', ml, n, 1.0, Aii, lda, Qii, ldq);
    if (prec == 0) return;
    size_t mark = ld_ws_mark();
    double *G = wsd((size_t)n * n), *R = wsd((size_t)n * n);
    dv_gram(1, ml, n, Qii, ldq, G, n);
    ld_potrf_u(n, G, n);
    ld_trsm('R', 'U', 'N', 'N', ml, n, 1.0, G, n, Qii, ldq);
    gemm(N_, N_, n, n, n, 1.0, G, n, LD_UPPER, Aii, lda, LD_UPPER, 0.0, R, n);
    ld_lacpy('U', n, n, R, n, Aii, lda);
    ld_ws_release(mark);
}

/* Cholesky-QR and the Householder reconstruction of one block (the whole
 * n columns inside one T block, so the reconstruction has no earlier blocks:
 * lila-c/lila_dgeqrf_w03_mt_hr.c:18-29 with a single lila_ormhr2_w03_hr
 * step), arranged so that the ml x n data is read or written five times
 * (one Gram pass) instead of about fourteen:
 *   G = A^T A = R^T R (A read in place, no copy of A into Q first);
 *   Q_top = A_top R^-1 (n x n) and its sign-chosen LU: Q_top S - I = L U;
 *   Q = A (R^-1 S), one GEMM with an upper triangular right operand, which
 *     is the reference's Q = A R^-1 with its columns already flipped;
 *   V2 = (Q S)_2 U^-1 into A below the top block, one more such GEMM;
 *   R <- S R, V1 = L, T = -U V1^-T on the n x n top block.
 * The same V, R, T and Q as the TRSM-and-copy sequence, to rounding (the
 * triangular solves run as products with explicit inverses).
 * Cholesky-QR2 (prec 1, 2, 3: the first Gram in FP64, FP32 MFMA or split
 * BF16, the second in FP64) forms W = A R1^-1 in a workspace first and runs
 * the same steps on W with its factor R2, R = R2 R1: seven passes. */
static void cholqr_reconstruct_oneblock(int prec, int ml, int n, double *Aii, int lda, double *Tq,
                                        int ldt, double *Qii, int ldq, double *S) {
    size_t mark = ld_ws_mark();
    double *G = wsd((size_t)n * n), *Ri = wsd((size_t)n * n), *Ui = wsd((size_t)n * n);
    const double *X = Aii;  /* the matrix whose Cholesky-QR gives Q */
    int ldx = lda;
    double *R1 = NULL;
    if (prec != 0) {
        /* first pass: W = A R1^-1 (R1^-1 from a triangular solve on the
         * identity: exact zeros below its diagonal) */
        R1 = wsd((size_t)n * n);
        const int ldw = ml + (ml & 1);  /* even: the direct-to-LDS GEMM takes it */
        double *W = wsd((size_t)ldw * n);
        dv_gram(prec, ml, n, Aii, lda, R1, n);
        ld_potrf_u(n, R1, n);
        ld_laset('A', n, n, 0.0, 1.0, Ri, n);
        ld_trsm('L', 'U', 'N', 'N', n, n, 1.0, R1, n, Ri, n);
        gemm(N_, N_, ml, n, n, 1.0, Aii, lda, LD_GEN, Ri, n, LD_UPPER_Z, 0.0, W, ldw);
        X = W;
        ldx = ldw;
    }
    dv_gram(prec != 0 ? 1 : 0, ml, n, X, ldx, G, n);
    ld_potrf_u(n, G, n);
    /* R^-1 (upper) */
    ld_laset('A', n, n, 0.0, 1.0, Ri, n);
    ld_trsm('L', 'U', 'N', 'N', n, n, 1.0, G, n, Ri, n);
    /* Q_top = X_top R^-1 into the T block, its sign-chosen LU there */
    gemm(N_, N_, n, n, n, 1.0, X, ldx, LD_GEN, Ri, n, LD_UPPER, 0.0, Tq, ldt);
    ld_lu_sign(n, Tq, ldt, S);
    /* Q = X (R^-1 S) */
    ld_flip_cols(n, n, Ri, n, S);
    /* (R^-1 and U^-1 come from triangular solves on the identity: exact zeros
     * below their diagonals, LD_UPPER_Z) */
    gemm(N_, N_, ml, n, n, 1.0, X, ldx, LD_GEN, Ri, n, LD_UPPER_Z, 0.0, Qii, ldq);
    /* V2 = (Q S)_2 U^-1 */
    ld_laset('A', n, n, 0.0, 1.0, Ui, n);
    ld_trsm('L', 'U', 'N', 'N', n, n, 1.0, Tq, ldt, Ui, n);
    if (ml > n) gemm(N_, N_, ml - n, n, n, 1.0, Qii + n, ldq, LD_GEN, Ui, n, LD_UPPER_Z, 0.0, Aii + n, lda);
    /* top block: R <- S R (R = R2 R1 after two passes), V1 = L, T = -U V1^-T */
    if (R1) {
        gemm(N_, N_, n, n, n, 1.0, G, n, LD_UPPER, R1, n, LD_UPPER, 0.0, Ri, n);
        ld_lacpy('U', n, n, Ri, n, Aii, lda);
    } else {
        ld_lacpy('U', n, n, G, n, Aii, lda);
    }
    ld_flip_rows_upper(n, Aii, lda, S);
    ld_move_lower(n, Tq, ldt, Aii, lda);
    ld_trsm('R', 'L', 'T', 'U', n, n, -1.0, Aii, lda, Tq, ldt);
    ld_ws_release(mark);
}

/* CholQR leaf + blocked reconstruction; lila-c/lila_dgeqrf_w03_mt_hr.c:3-33 */
int dv_dgeqrf_w03_mt_hr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                        int ldq) {
    int ml = m - i;
    if (i % mt + n <= mt && ml >= n) {
        size_t mark = ld_ws_mark();
        double *Sd = wsd((size_t)(i + n));
        cholqr_reconstruct_oneblock(ld_tune()->gram_prec, ml, n, A + i + (size_t)i * lda, lda,
                                    T + i % mt + (size_t)i * ldt, ldt, Q + i + (size_t)i * ldq, ldq,
                                    Sd + i);
        ld_ws_release(mark);
        return 0;
    }
    cholqr(ml, n, A + i + (size_t)i * lda, lda, Q + i + (size_t)i * ldq, ldq);
    size_t mark = ld_ws_mark();
    double *Sd = wsd((size_t)(i + n));
    dv_dorghr(m, n, i, mt, A, lda, T, ldt, Q, ldq, Sd);
    ld_ws_release(mark);
    return 0;
}

/* CholQR panel + dorgh2_3; lila-c/lila_dgeqr2_w03_hr.c:3-30 */
int dv_dgeqr2_w03_hr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                     int ldq) {
    int ml = m - i;
    double *Aii = A + i + (size_t)i * lda, *Qii = Q + i + (size_t)i * ldq;
    double *Tki = T + i % mt + (size_t)i * ldt;
    if (ml >= n) {
        size_t mark = ld_ws_mark();
        double *S = wsd((size_t)n);
        cholqr_reconstruct_oneblock(ld_tune()->gram_prec, ml, n, Aii, lda, Tki, ldt, Qii, ldq, S);
        ld_ws_release(mark);
        return 0;
    }
    cholqr(ml, n, Aii, lda, Qii, ldq);
    ld_lacpy('A', ml - n, n, Qii + n, ldq, Aii + n, lda);
    ld_lacpy('A', n, n, Qii, ldq, Tki, ldt);
    size_t mark = ld_ws_mark();
    double *S = wsd((size_t)n);
    dv_dorgh2_3(ml, n, Aii, lda, Tki, ldt, Qii, ldq, S);
    ld_ws_release(mark);
    return 0;
}

/* panel dispatch on lila_param[2]; lila-c/lila_dgeqr2_w03.c:8-16 */
int dv_dgeqr2_w03(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                  double *Q, int ldq) {
    switch (p[2]) {
    case 0: return dv_dgeqr2_w03_l(m, n, i, mt, A, lda, T, ldt, Q, ldq);
    case 1: return dv_dgeqr2_w03_3(m, n, i, mt, A, lda, T, ldt, Q, ldq);
    case 2: return dv_dgeqr2_w03_hr(m, n, i, mt, A, lda, T, ldt, Q, ldq);
    }
    return 0;
}

/* per-mt-block T from V and tau (lila-c/lila_dlarft_w03.c:3-55) */
int dv_dlarft_w03(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, const double *tau) {
    int j = i, vb = mt - i % mt;
    if (vb > n) vb = n;
    while (vb > 0) {
        ld_larft(m - j, vb, A + j + (size_t)j * lda, lda, tau + (j - i), T + j % mt + (size_t)j * ldt, ldt);
        j += vb;
        vb = (j + mt >= i + n) ? (i + n - j) : mt;
    }
    return 0;
}

/* LAPACK leaf (dgeqrf, lila_dlarft_w03, dorgqr; lila-c/lila_dgeqrf_w03_mt_l.c):
 * the recursive panel gives V and the full T (whose diagonal is tau); T is
 * then laid out per mt-block by dlarft_w03 and Q = [I;0] - V T V1^T. */
int dv_dgeqrf_w03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                       int ldq) {
    int ml = m - i;
    double *Aii = A + i + (size_t)i * lda, *Qii = Q + i + (size_t)i * ldq;
    size_t mark = ld_ws_mark();
    double *Tf = wsd((size_t)n * n), *tau = wsd((size_t)n);
    ld_laset('L', n, n, 0.0, 0.0, Tf, n);
    dv_dgeqr3(ml, n, Aii, lda, Tf, n);
    ld_T2tau(n, 0, n + 1, Tf, n, tau);
    dv_dlarft_w03(m, n, i, mt, A, lda, T, ldt, tau);
    panel_q(ml, n, Aii, lda, Tf, n, Qii, ldq);
    ld_ws_release(mark);
    return 0;
}

/* hh leaf; lila-c/lila_dgeqrf_w03_mt_hh.c:7-32 */
int dv_dgeqrf_w03_mt_hh(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                        int ldt, double *Q, int ldq) {
    int vb = mt - i % mt;
    if (vb > n) vb = n;
    dv_dgeqr2_w03(p, m, vb, i, mt, A, lda, T, ldt, Q, ldq);
    int j = i + vb, l = vb;
    vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
    while (vb != 0) {
        dv_dormqrf_w03(m, vb, l, i, j, mt, A, lda, T, ldt);
        dv_dgeqr2_w03(p, m, vb, j, mt, A, lda, T, ldt, Q, ldq);
        dv_dormqrbz_w03(m, vb, l, i, j, mt, A, lda, Q, ldq, T, ldt);
        j += vb;
        l += vb;
        vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
    }
    return 0;
}

/* leaf dispatch on lila_param[1]; lila-c/lila_dgeqrf_w03_mt.c:6-16 */
int dv_dgeqrf_w03_mt(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                     double *Q, int ldq) {
    switch (p[1]) {
    case 0: return dv_dgeqrf_w03_mt_l(m, n, i, mt, A, lda, T, ldt, Q, ldq);
    case 1: return dv_dgeqrf_w03_mt_hh(p, m, n, i, mt, A, lda, T, ldt, Q, ldq);
    case 2: return dv_dgeqrf_w03_mt_hr(m, n, i, mt, A, lda, T, ldt, Q, ldq);
    }
    return 0;
}

/* Mixed precision (lila_param[0] == 1): the R0 of a Householder QR whose
 * trailing updates and T products run on FP32 MFMA (panels in FP64)
 * preconditions A; FP64 Cholesky-QR twice on A R0^-1 gives an orthonormal Q
 * and R = R2 R1 R0; the blocked Householder reconstruction (lila_dorghr)
 * turns (Q, R) into the same V, R, T and Q outputs as the FP64 path, in
 * exact arithmetic and to FP64 accuracy for kappa(A) up to ~1/(u32 c). */
int dv_dgeqrf_mixed(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                    double *Q, int ldq) {
    int ml = m - i;
    if (n <= 0) return 0;
    double *Aii = A + i + (size_t)i * lda, *Qii = Q + i + (size_t)i * ldq;
    size_t mark = ld_ws_mark();
    double *W = wsd((size_t)ml * n), *T0 = wsd((size_t)mt * n);
    double *G = wsd((size_t)n * n), *R = wsd((size_t)n * n), *R2 = wsd((size_t)n * n);
    double *Sd = wsd((size_t)(i + n));
    /* R0: FP32-MFMA updates, FP64 panels */
    ld_lacpy('A', ml, n, Aii, lda, W, ml);
    /* leaves of at most nx columns, as in the FP64 recursion (the VRT
     * recursion stops at n < lila_param[3], the VRTQ one at n <= it) */
    int p0[4] = {0, 1, 1, (p[3] > 1 ? p[3] : 32) + 1};
    ld_tune_t *tu = ld_tune();
    const int old = tu->gemm_lowp;
    tu->gemm_lowp = 1;
    dv_dgeqrf_v03_recursive(p0, ml, n, 0, mt, W, ml, T0, mt);
    tu->gemm_lowp = old;
    /* Q1 = A R0^-1 and R = R0 */
    ld_lacpy('A', ml, n, Aii, lda, Qii, ldq);
    ld_trsm('R', 'U', 'N', 'N', ml, n, 1.0, W, ml, Qii, ldq);
    ld_laset('L', n, n, 0.0, 0.0, R, n);
    ld_lacpy('U', n, n, W, ml, R, n);
    /* Cholesky-QR: G = Q^T Q = Rc^T Rc, Q <- Q Rc^-1, R <- Rc R.  A second
     * pass only when the first Gram shows Q1 away from orthonormal:
     * ||I - G||_2 <= 0.5 bounds kappa(Q1)^2 by 3, and one pass then leaves
     * ||I - Q^T Q||_F at ~0.4 n u, the Householder path's level.  First the
     * exact bound ||I - G||_2 <= ||I - G||_F <= 0.5; failing that, unless
     * ||I - G||_F / sqrt(n) (<= ||I - G||_2) already exceeds 1/6, a power-
     * method estimate, which is a lower bound and so must stay below 1/6
     * (a factor 3 margin).  At m = n = 16384 with a FP32-updated R0:
     * uniform entries give ||I - G||_F = 0.16; kappa(A) = 1e5 gives 1.1 and
     * ||I - G||_2 = 0.042 (one pass); kappa(A) = 1e6 gives 0.44 and 1e8
     * gives 3.2e6 (two passes). */
    int npass = 2;
    for (int pass = 0; pass < npass; pass++) {
        ld_syrk_ut(n, ml, 1.0, Qii, ldq, 0.0, G, n);
        if (pass == 0) {
            const double df = ld_orth_defect(n, G, n);
            if (df <= 0.5 ||
                (df <= sqrt((double)n) / 6.0 && 3.0 * ld_sym_defect_norm2_est(n, G, n, 20) <= 0.5))
                npass = 1;
        }
        ld_potrf_u(n, G, n);
        ld_trsm('R', 'U', 'N', 'N', ml, n, 1.0, G, n, Qii, ldq);
        /* R <- Rc R: both upper triangular, so the GEMM's k range per tile
         * covers only their overlap */
        gemm(N_, N_, n, n, n, 1.0, G, n, LD_UPPER, R, n, LD_UPPER, 0.0, R2, n);
        double *t = R;
        R = R2;
        R2 = t;
    }
    /* A (rank deficient, or kappa far beyond 1e8) that makes the low-
     * precision R0 singular or a Cholesky factor break down is factored by
     * the FP64 recursion instead, like the reference would: A is untouched
     * up to here, and T and Q are outputs only */
    if (ld_check_device_flags() > 0) {
        ld_clear_info();
        ld_ws_release(mark);
        int p64[4] = {0, p[1], p[2], p[3]};
        return dv_dgeqrf_w03_recursive(p64, m, n, i, mt, A, lda, T, ldt, Q, ldq);
    }
    /* Householder reconstruction of (Q, R) in the mt layout */
    ld_lacpy('U', n, n, R, n, Aii, lda);
    dv_dorghr(m, n, i, mt, A, lda, T, ldt, Q, ldq, Sd);
    ld_ws_release(mark);
    return 0;
}

/* lila-c/lila_dgeqrf_w03_recursive.c:3-29; lila_param[0] == 1 selects the
 * mixed-precision form above */
/* The Q-building step of a node (dormqrbz) reads the node's left reflectors
 * (V below the diagonal of columns [i, i+n1), their T blocks) and the right
 * half's Q columns, and writes Q's columns [i+n1, i+n).  Nothing the
 * factorization does afterwards touches those: it writes A and T only in
 * columns further right (the parent's connect writes T rows above the
 * diagonal blocks read here) and Q only through its own later leaves and
 * Q-building steps.  So the Q-building steps run on lane 1, each after a
 * fork that orders it behind the node's factorization on lane 0, and the
 * outermost call joins lane 1 back before it returns. */
static int w03_depth = 0;

static int w03_rec(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                   double *Q, int ldq) {
    if (n <= p[3]) return dv_dgeqrf_w03_mt(p, m, n, i, mt, A, lda, T, ldt, Q, ldq);
    int n1 = n / 2, n2 = n - n1;
    w03_rec(p, m, n1, i, mt, A, lda, T, ldt, Q, ldq);
    dv_dormqrf_w03(m, n2, n1, i, i + n1, mt, A, lda, T, ldt);
    w03_rec(p, m, n2, i + n1, mt, A, lda, T, ldt, Q, ldq);
    dv_dlarft_connect_w03(m, n2, i + n1, i, mt, A, lda, T, ldt);
    if (ld_lanes() > 1) {
        ld_lane_fork();
        ld_set_lane(1);
        dv_dormqrbz_w03(m, n2, n1, i, i + n1, mt, A, lda, Q, ldq, T, ldt);
        ld_set_lane(0);
    } else {
        dv_dormqrbz_w03(m, n2, n1, i, i + n1, mt, A, lda, Q, ldq, T, ldt);
    }
    return 0;
}

/* Q(i:m, i:i+n) = H_i ... H_{i+n-1} [I; 0] from V (below A's diagonal) and
 * the mt-layout T, by T blocks from the last to the first (LAPACK dorgqr's
 * order): block b's reflectors applied to the columns right of it, which
 * hold [0; Q] below the block (lila_dormqrbz_w03's step), and block b's own
 * columns [I; 0] - V_b T_b V_b1^T.  Every block step is two or three GEMMs
 * as wide as the columns right of the block. */
static void q_from_vt(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                      int ldq) {
    if (n <= 0) return;
    const int nblk = (n + mt - 1) / mt + 1;
    int *b0 = (int *)malloc(sizeof(int) * (size_t)nblk), *wb = (int *)malloc(sizeof(int) * (size_t)nblk);
    if (!b0 || !wb) {
        free(b0); free(wb);
        ld_set_info(-2);
        return;
    }
    int nb = 0, j = i, w = mt - i % mt;
    if (w > n) w = n;
    while (w > 0) {
        b0[nb] = j;
        wb[nb++] = w;
        j += w;
        w = (j + mt >= i + n) ? (i + n - j) : mt;
    }
    for (int b = nb - 1; b >= 0; b--) {
        const int c0 = b0[b], cw = wb[b], r0 = c0 + cw, nr = i + n - r0;
        if (nr > 0) dv_dormqrbz_w03(m, nr, cw, c0, r0, mt, A, lda, Q, ldq, T, ldt);
        panel_q(m - c0, cw, A + c0 + (size_t)c0 * lda, lda, T + c0 % mt + (size_t)c0 * ldt, ldt,
                Q + c0 + (size_t)c0 * ldq, ldq);
    }
    free(b0);
    free(wb);
}

int dv_dgeqrf_w03_recursive(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                            int ldt, double *Q, int ldq) {
    if (p[0] == 1) return dv_dgeqrf_mixed(p, m, n, i, mt, A, lda, T, ldt, Q, ldq);
    /* The VRT recursion with the reference's leaves (n <= nx) gives the same
     * V, R and T as the VRTQ call tree; building Q afterwards by whole T
     * blocks replaces the Q steps at every node and leaf (many narrow
     * products near the bottom of the recursion) with a few wide ones.  Q is
     * the same matrix (H_i ... H_{i+n-1} [I; 0]) to rounding. */
    if (ld_tune()->q_after && ld_lanes() == 1 && (p[1] == 0 || p[1] == 1) && (p[2] == 0 || p[2] == 1)) {
        const int pv[4] = {0, p[1], p[2], p[3] + 1};
        dv_dgeqrf_v03_recursive(pv, m, n, i, mt, A, lda, T, ldt);
        q_from_vt(m, n, i, mt, A, lda, T, ldt, Q, ldq);
        return 0;
    }
    w03_depth++;
    w03_rec(p, m, n, i, mt, A, lda, T, ldt, Q, ldq);
    if (--w03_depth == 0 && ld_lanes() > 1) ld_lane_join();
    return 0;
}

/* lila-c/lila_dgeqrf_w03_levelx.c:3-52 (i_lvl by value) */
int dv_dgeqrf_w03_levelx(const int *p, int n_lvl, int i_lvl, const int *nb_lvl, int m, int n, int i,
                         int mt, double *A, int lda, double *T, int ldt, double *Q, int ldq) {
    int nb = nb_lvl[i_lvl];
    int vb = nb > n ? n : nb;
    int k = 0, j = i;
    if (i_lvl == n_lvl - 1) dv_dgeqrf_w03_mt(p, m, vb, i, mt, A, lda, T, ldt, Q, ldq);
    else dv_dgeqrf_w03_levelx(p, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt, A, lda, T, ldt, Q, ldq);
    k += vb;
    j += vb;
    while (j < i + n) {
        vb = (j + nb > i + n) ? (i + n - j) : nb;
        dv_dormqrf_w03(m, vb, k, i, j, mt, A, lda, T, ldt);
        if (i_lvl == n_lvl - 1) dv_dgeqrf_w03_mt(p, m, vb, j, mt, A, lda, T, ldt, Q, ldq);
        else dv_dgeqrf_w03_levelx(p, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt, A, lda, T, ldt, Q, ldq);
        dv_dlarft_connect_w03(m, vb, j, i, mt, A, lda, T, ldt);
        dv_dormqrbz_w03(m, vb, k, i, j, mt, A, lda, Q, ldq, T, ldt);
        k += vb;
        j += vb;
    }
    return 0;
}

/* c3/lila_dgetrf_b.c:3-34; U (R, n x n ld n) is the device work array */
int dv_dgetrf_b(int m, int n, int i, int mt, double *A, int lda, double *Q, int ldq, double *U) {
    (void)mt;
    ld_getrf_b(m - i, n, A + i + (size_t)i * lda, lda, Q + i + (size_t)i * ldq, ldq, U);
    return 0;
}

/* ---- the recursive QR library of qr3/src (qr2.h) -------------------------- */

/* qr3/src/qr2_dgeqr3_R.c:3-60: the recursive panel with R kept apart: V
 * below A's diagonal, T (n x n upper), R (n x n upper).  A's diagonal and
 * upper part are no output there (they end as each entry stood before its
 * column's reflector); here they keep their input values.  With R == A the
 * result is dgeqr3's. */
int dv_qr2_dgeqr3_R(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr) {
    if (n <= 0) return 0;
    if (R == A && ldr == lda) return dv_dgeqr3(m, n, A, lda, T, ldt);
    size_t mark = ld_ws_mark();
    double *U = wsd((size_t)n * n);
    ld_lacpy('U', n, n, A, lda, U, n);
    dv_dgeqr3(m, n, A, lda, T, ldt);
    ld_lacpy('U', n, n, A, lda, R, ldr);
    ld_lacpy('U', n, n, U, n, A, lda);
    ld_ws_release(mark);
    return 0;
}

/* qr3/src/qr2_dlarft3.c:3-44: T (n x n upper, dlarft 'F','C') from the
 * reflectors below A's diagonal and tau */
int dv_qr2_dlarft3(int m, int n, double *A, int lda, double *T, int ldt, double *tau) {
    if (n <= 0) return 0;
    ld_larft(m, n, A, lda, tau, T, ldt);
    return 0;
}

/* qr3/src/qr2_dorgqr3.c:3-58: the explicit m x n Q = H_1 ... H_n [I; 0]
 * over the reflectors stored below Q's diagonal, in place: Q = [I; 0] -
 * V (T V1^T) with V copied out first.  T's diagonal stands for tau (the
 * reference reads tau at its 1-column leaves, T elsewhere; they agree for
 * any T of these reflectors). */
int dv_qr2_dorgqr3(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau) {
    (void)tau;
    if (n <= 0) return 0;
    size_t mark = ld_ws_mark();
    double *V = wsd((size_t)m * n);
    ld_lacpy('L', m, n, Q, ldq, V, m);
    panel_q(m, n, V, m, T, ldt, Q, ldq);
    ld_ws_release(mark);
    return 0;
}

/* ---- the UT, ISW and VT2Q/VS2Q forms of qr3/src and its aux routines ----
 *
 * UT form: the block reflector I - V S^-1 V^T is carried by S = T^-1, whose
 * strict upper part is that of V^T V and whose diagonal is 1/tau (the
 * reference's UT variants build exactly that).  ISW forms: the reference's
 * recursion computes full T (or S) blocks only for the left halves down its
 * right spine -- diagonal blocks of sizes n/2, (n - n/2)/2, ..., 1 -- and
 * never writes the off-diagonal blocks between them; here the full T (or S)
 * is formed in the workspace and only those diagonal blocks are copied out. */

static void c_upper_gemm(int ta, int tb, int n, int k, double alpha, const double *A, int lda, int sa,
                         const double *B, int ldb, int sb, double *C, int ldc) {
    ld_gemm_t g;
    g.m = n; g.n = n; g.k = k; g.alpha = alpha; g.beta = 0.0;
    g.A = A; g.lda = lda; g.ta = ta; g.sa = sa;
    g.B = B; g.ldb = ldb; g.tb = tb; g.sb = sb;
    g.C = C; g.ldc = ldc; g.zero_rows = 0; g.c_upper = 1;
    ld_gemm(&g);
}

/* S(upper) = triu(V^T V) with diagonal 1/tau (tau[c * tinc]) */
static void ut_from_v(int m, int n, const double *V, int ldv, const double *tau, int tinc, double *S,
                      int lds) {
    c_upper_gemm(T_, N_, n, m, 1.0, V, ldv, LD_VUNIT, V, ldv, LD_VUNIT, S, lds);
    ld_set_diag(n, tau, tinc, 1, S, lds);
}

/* the diagonal blocks of the right spine of the halving recursion */
static void copy_spine(int n, const double *W, int ldw, double *T, int ldt) {
    int off = 0;
    while (n > 1) {
        const int n1 = n / 2;
        ld_lacpy('U', n1, n1, W + off + (size_t)off * ldw, ldw, T + off + (size_t)off * ldt, ldt);
        off += n1;
        n -= n1;
    }
    if (n == 1) ld_lacpy('U', 1, 1, W + off + (size_t)off * ldw, ldw, T + off + (size_t)off * ldt, ldt);
}

/* X (n x n) = S^-1 for the upper S */
static void inv_upper(int n, const double *S, int lds, double *X) {
    ld_laset('A', n, n, 0.0, 1.0, X, n);
    ld_trsm('L', 'U', 'N', 'N', n, n, 1.0, S, lds, X, n);
}

/* qr3/src/qr2_dgeqr3_R_UT.c:3-67: V and R as qr2_dgeqr3_R, T <- S */
int dv_qr2_dgeqr3_R_UT(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr) {
    if (n <= 0) return 0;
    dv_qr2_dgeqr3_R(m, n, A, lda, T, ldt, R, ldr);
    size_t mark = ld_ws_mark();
    double *tau = wsd((size_t)n);
    ld_lacpy('A', 1, n, T, ldt + 1, tau, 1);
    ut_from_v(m, n, A, lda, tau, 1, T, ldt);
    ld_ws_release(mark);
    return 0;
}

/* qr3/src/qr2_dgeqr3_R_ISW.c:3-63 and qr2_dgeqr3_R_UT_ISW.c:3-64: V and R
 * as qr2_dgeqr3_R; T (or S) on the spine blocks only */
int dv_qr2_dgeqr3_R_isw(int ut, int m, int n, double *A, int lda, double *T, int ldt, double *R,
                        int ldr) {
    if (n <= 0) return 0;
    size_t mark = ld_ws_mark();
    double *Tw = wsd((size_t)n * n);
    if (ut) dv_qr2_dgeqr3_R_UT(m, n, A, lda, Tw, n, R, ldr);
    else dv_qr2_dgeqr3_R(m, n, A, lda, Tw, n, R, ldr);
    copy_spine(n, Tw, n, T, ldt);
    ld_ws_release(mark);
    return 0;
}

/* qr3/src/qr2_dlarft3_UT.c:3-43 */
int dv_qr2_dlarft3_UT(int m, int n, double *A, int lda, double *T, int ldt, double *tau) {
    if (n <= 0) return 0;
    ut_from_v(m, n, A, lda, tau, 1, T, ldt);
    return 0;
}

/* qr3/src/qr2_dlarft3_ISW.c:3-39, qr2_dlarft3_ISW_V2T.c:3-43 (ut = 0) and
 * qr2_dlarft3_UT_ISW.c:3-39, qr2_dlarft3_ISW_V2T_UT.c:3-43 (ut = 1): the
 * V2T forms reach the same spine blocks through V^T V */
int dv_qr2_dlarft3_isw(int ut, int m, int n, double *A, int lda, double *T, int ldt, double *tau) {
    if (n <= 0) return 0;
    size_t mark = ld_ws_mark();
    double *Tw = wsd((size_t)n * n);
    if (ut) ut_from_v(m, n, A, lda, tau, 1, Tw, n);
    else ld_larft(m, n, A, lda, tau, Tw, n);
    copy_spine(n, Tw, n, T, ldt);
    ld_ws_release(mark);
    return 0;
}

/* qr3/src/qr2_dorgqr3_UT.c:3-53 and qr2_dorgqr3_VS2Q.c:3-54: Q from V (below
 * Q's diagonal) and S = T^-1 */
int dv_qr2_dorgqr3_UT(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau) {
    (void)tau;
    if (n <= 0) return 0;
    size_t mark = ld_ws_mark();
    double *X = wsd((size_t)n * n), *V = wsd((size_t)m * n);
    inv_upper(n, T, ldt, X);
    ld_lacpy('L', m, n, Q, ldq, V, m);
    panel_q(m, n, V, m, X, n, Q, ldq);
    ld_ws_release(mark);
    return 0;
}

/* qr3/src/qr2_aux_dVT2Q.c:3-17: Q (m x n) holds V below its diagonal and T
 * on and above; Q <- [I;0] - V T V1^T */
int dv_qr2_aux_dVT2Q(int m, int n, double *Q, int ldq) {
    if (n <= 0) return 0;
    size_t mark = ld_ws_mark();
    double *W = wsd((size_t)m * n);
    ld_lacpy('A', m, n, Q, ldq, W, m);
    panel_q(m, n, W, m, W, m, Q, ldq);
    ld_ws_release(mark);
    return 0;
}

/* qr3/src/qr2_aux_dVS2Q.c:3-17: the same with S = T^-1 on and above */
int dv_qr2_aux_dVS2Q(int m, int n, double *Q, int ldq) {
    if (n <= 0) return 0;
    size_t mark = ld_ws_mark();
    double *X = wsd((size_t)n * n), *V = wsd((size_t)m * n);
    inv_upper(n, Q, ldq, X);
    ld_lacpy('L', m, n, Q, ldq, V, m);
    panel_q(m, n, V, m, X, n, Q, ldq);
    ld_ws_release(mark);
    return 0;
}

/* qr3/src/qr2_aux_dmLUinA.c:3-40: A <- -L U (L unit lower, U upper, both in A) */
int dv_qr2_aux_dmLUinA(int n, double *A, int lda) {
    if (n <= 0) return 0;
    size_t mark = ld_ws_mark();
    double *W = wsd((size_t)n * n);
    ld_lacpy('A', n, n, A, lda, W, n);
    gemm(N_, N_, n, n, n, -1.0, W, n, LD_VUNIT, W, n, LD_UPPER, 0.0, A, lda);
    ld_ws_release(mark);
    return 0;
}

/* qr3/src/qr2_aux_dULTinU.c:3-38: upper(U) <- U L^T (L unit lower; U's strict
 * lower part is not touched, so L may share U's storage) */
int dv_qr2_aux_dULTinU(int n, double *L, int ldl, double *U, int ldu) {
    if (n <= 0) return 0;
    size_t mark = ld_ws_mark();
    double *W = wsd((size_t)n * n);
    ld_lacpy('A', n, n, U, ldu, W, n);
    const int alias = (L == U && ldl == ldu);
    c_upper_gemm(N_, T_, n, n, 1.0, W, n, LD_UPPER, alias ? W : L, alias ? n : ldl, LD_VUNIT, U, ldu);
    ld_ws_release(mark);
    return 0;
}

/* qr3/src/qr2_aux_dUinvLTinU.c:3-37: upper(U) <- U^-1 L^T */
int dv_qr2_aux_dUinvLTinU(int n, double *L, int ldl, double *U, int ldu) {
    if (n <= 0) return 0;
    size_t mark = ld_ws_mark();
    double *X = wsd((size_t)n * n);
    ld_laset('A', n, n, 0.0, 1.0, X, n);
    ld_trmm('R', 'L', 'T', 'U', n, n, 1.0, L, ldl, X, n);
    ld_trsm('L', 'U', 'N', 'N', n, n, 1.0, U, ldu, X, n);
    ld_lacpy('U', n, n, X, n, U, ldu);
    ld_ws_release(mark);
    return 0;
}

/* qr3/src/qr2_aux_dApUBTinA.c:46-96: A (m x n) += U B^T, U (m x m) upper,
 * B n x m */
int dv_qr2_aux_dApUBTinA(int m, int n, double *A, int lda, double *U, int ldu, double *B, int ldb) {
    if (m <= 0 || n <= 0) return 0;
    gemm(N_, T_, m, n, m, 1.0, U, ldu, LD_UPPER, B, ldb, LD_GEN, 1.0, A, lda);
    return 0;
}

/* qr3/src/qr2_aux_dV2N.c:3-30: upper(T) <- U U^T for the unit upper U whose
 * strict upper part is T's (for T holding V1^T: V1^T V1) */
int dv_qr2_aux_dV2N(int n, double *T, int ldt) {
    if (n <= 0) return 0;
    size_t mark = ld_ws_mark();
    double *W = wsd((size_t)n * n);
    ld_lacpy('U', n, n, T, ldt, W, n);
    c_upper_gemm(N_, T_, n, n, 1.0, W, n, LD_UUNIT, W, n, LD_UUNIT, T, ldt);
    ld_ws_release(mark);
    return 0;
}

/* qr3/src/qr2_aux_dN2T.c:3-40 */
int dv_qr2_aux_dN2T(int n, double *tau, double *T, int ldt) {
    ld_n2t(n, tau, T, ldt);
    return 0;
}

/* qr3/src/qr2_aux_dV2tau.c:3-15 */
int dv_qr2_aux_dV2tau(int m, int n, double *A, int lda, double *tau) {
    if (n <= 0) return 0;
    ld_V2tau(m, n, 0, A, lda, tau);
    if (m == n) ld_laset('A', 1, 1, 0.0, 0.0, tau + n - 1, 1);
    return 0;
}

/* qr3/src/qr2_aux_dlarf_wrapper.c:3-8 (LAPACK dlarf): C <- H C (side 'L')
 * or C H ('R'), H = I - tau v v^T, v with stride incv >= 1 */
int dv_qr2_aux_dlarf(char side, int m, int n, const double *V, int incv, double tau, double *C,
                     int ldc) {
    if (m <= 0 || n <= 0 || tau == 0.0) return 0;
    const int len = (side == 'L') ? m : n;
    size_t mark = ld_ws_mark();
    double *v = wsd((size_t)len), *w = wsd((size_t)(side == 'L' ? n : m));
    ld_lacpy('A', 1, len, V, incv, v, 1);
    if (side == 'L') {
        gemm(T_, N_, 1, n, m, 1.0, v, m, LD_GEN, C, ldc, LD_GEN, 0.0, w, 1);
        gemm(N_, N_, m, n, 1, -tau, v, m, LD_GEN, w, 1, LD_GEN, 1.0, C, ldc);
    } else {
        gemm(N_, N_, m, 1, n, 1.0, C, ldc, LD_GEN, v, n, LD_GEN, 0.0, w, m);
        gemm(N_, T_, m, n, 1, -tau, w, m, LD_GEN, v, n, LD_GEN, 1.0, C, ldc);
    }
    ld_ws_release(mark);
    return 0;
}

/* least squares through (V, T, R): lila_dgels_qr3 semantic of
 * least-squares/main_ls_qr3.c:94-119.  B (m x k) overwritten, X in rows [0, n). */
int dv_ls_qr3(int m, int n, int k, double *A, int lda, double *T, int ldt, double *B, int ldb) {
    dv_dgeqr3(m, n, A, lda, T, ldt);
    apply_block_qt(m, k, n, A, lda, T, ldt, B, ldb);
    ld_trsm('L', 'U', 'N', 'N', n, k, 1.0, A, lda, B, ldb);
    return 0;
}

/* ---- VRT only (v03), Q from (V, T) (q03), append-columns, LAPACK-kernel
 * and qr2 families (lila-c/lila.h:19-22, :74-90, :95-102) ------------------ */

/* C <- (I - V T V^T) C for one block (the non-transposed application) */
static void apply_block_q(int ml, int n, int vb, const double *V, int ldv, const double *Tb, int ldt,
                          double *C, int ldc) {
    if (vb <= 0 || n <= 0) return;
    size_t mark = ld_ws_mark();
    double *W1 = wsd((size_t)vb * n), *W2 = wsd((size_t)vb * n);
    gemm(T_, N_, vb, n, ml, 1.0, V, ldv, LD_VUNIT, C, ldc, LD_GEN, 0.0, W1, vb);
    gemm(N_, N_, vb, n, vb, 1.0, Tb, ldt, LD_UPPER, W1, vb, LD_GEN, 0.0, W2, vb);
    gemm(N_, N_, ml, n, vb, -1.0, V, ldv, LD_VUNIT, W2, vb, LD_GEN, 1.0, C, ldc);
    ld_ws_release(mark);
}

/* LAPACK dgeqrf on the GPU: the recursive panel's V and R, tau = diag of its
 * T (dgeqr3 uses dlarfg, so these are dgeqrf's V, R and tau) */
static void geqrf_tau(int ml, int n, double *A, int lda, double *tau) {
    size_t mark = ld_ws_mark();
    double *Tf = wsd((size_t)n * n);
    dv_dgeqr3(ml, n, A, lda, Tf, n);
    ld_T2tau(n, 0, n + 1, Tf, n, tau);
    ld_ws_release(mark);
}

/* LAPACK dormqr('L', trans) with k reflectors given by V and tau: one block
 * reflector with T = dlarft(V, tau) */
static void orm_tau(char trans, int ml, int n, int k, const double *V, int ldv, const double *tau,
                    double *C, int ldc) {
    if (k <= 0 || n <= 0) return;
    size_t mark = ld_ws_mark();
    double *Tf = wsd((size_t)k * k);
    ld_larft(ml, k, V, ldv, tau, Tf, k);
    if (trans == 'T') apply_block_qt(ml, n, k, V, ldv, Tf, k, C, ldc);
    else apply_block_q(ml, n, k, V, ldv, Tf, k, C, ldc);
    ld_ws_release(mark);
}

/* LAPACK dorgqr: the first n columns of H_0 ... H_{n-1} */
static void org_tau(int ml, int n, const double *V, int ldv, const double *tau, double *Q, int ldq) {
    size_t mark = ld_ws_mark();
    double *Tf = wsd((size_t)n * n);
    ld_larft(ml, n, V, ldv, tau, Tf, n);
    panel_q(ml, n, V, ldv, Tf, n, Q, ldq);
    ld_ws_release(mark);
}

/* lila-c/lila_dgeqr2_v03_3.c:3-31 (its T V1^T product is never read) and
 * lila-c/lila_dgeqr2_v03_l.c:3-25 (dgeqrf + dlarft: the same V, R, T) */
int dv_dgeqr2_v03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt) {
    return dv_dgeqr3(m - i, n, A + i + (size_t)i * lda, lda, T + i % mt + (size_t)i * ldt, ldt);
}

/* lila-c/lila_dgeqr2_v03.c:3-19 */
int dv_dgeqr2_v03(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt) {
    if (p[2] == 0 || p[2] == 1) return dv_dgeqr2_v03_3(m, n, i, mt, A, lda, T, ldt);
    return 0;
}

/* lila-c/lila_dgeqrf_v03_mt_hh.c:3-35 */
int dv_dgeqrf_v03_mt_hh(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                        int ldt) {
    int vb = mt - i % mt;
    if (vb > n) vb = n;
    dv_dgeqr2_v03(p, m, vb, i, mt, A, lda, T, ldt);
    int j = i + vb, l = vb;
    vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
    while (vb != 0) {
        dv_dormqrf_w03(m, vb, l, i, j, mt, A, lda, T, ldt);
        dv_dgeqr2_v03(p, m, vb, j, mt, A, lda, T, ldt);
        j += vb;
        l += vb;
        vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
    }
    return 0;
}

/* lila-c/lila_dgeqrf_v03_mt_l.c:3-20: dgeqrf, then lila_dlarft_w03 */
int dv_dgeqrf_v03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt) {
    size_t mark = ld_ws_mark();
    double *tau = wsd((size_t)n);
    geqrf_tau(m - i, n, A + i + (size_t)i * lda, lda, tau);
    dv_dlarft_w03(m, n, i, mt, A, lda, T, ldt, tau);
    ld_ws_release(mark);
    return 0;
}

/* lila-c/lila_dgeqrf_v03_mt.c:3-19 */
int dv_dgeqrf_v03_mt(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                     int ldt) {
    if (p[1] == 0) return dv_dgeqrf_v03_mt_l(m, n, i, mt, A, lda, T, ldt);
    if (p[1] == 1) return dv_dgeqrf_v03_mt_hh(p, m, n, i, mt, A, lda, T, ldt);
    return 0;
}

/* lila-c/lila_dgeqrf_v03_recursive.c:3-29 (leaf when n < nx) */
int dv_dgeqrf_v03_recursive(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                            int ldt) {
    if (n < p[3]) return dv_dgeqrf_v03_mt(p, m, n, i, mt, A, lda, T, ldt);
    int n1 = n / 2, n2 = n - n1;
    dv_dgeqrf_v03_recursive(p, m, n1, i, mt, A, lda, T, ldt);
    dv_dormqrf_w03(m, n2, n1, i, i + n1, mt, A, lda, T, ldt);
    dv_dgeqrf_v03_recursive(p, m, n2, i + n1, mt, A, lda, T, ldt);
    dv_dlarft_connect_w03(m, n2, i + n1, i, mt, A, lda, T, ldt);
    return 0;
}

/* lila-c/lila_dgeqrf_v03_levelx.c:3-52 */
int dv_dgeqrf_v03_levelx(const int *p, int n_lvl, int i_lvl, const int *nb_lvl, int m, int n, int i,
                         int mt, double *A, int lda, double *T, int ldt) {
    int nb = nb_lvl[i_lvl];
    int vb = nb > n ? n : nb;
    int k = 0, j = i;
    if (i_lvl == n_lvl - 1) dv_dgeqrf_v03_mt(p, m, vb, i, mt, A, lda, T, ldt);
    else dv_dgeqrf_v03_levelx(p, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt, A, lda, T, ldt);
    k += vb;
    j += vb;
    while (j < i + n) {
        vb = (j + nb > i + n) ? (i + n - j) : nb;
        dv_dormqrf_w03(m, vb, k, i, j, mt, A, lda, T, ldt);
        if (i_lvl == n_lvl - 1) dv_dgeqrf_v03_mt(p, m, vb, j, mt, A, lda, T, ldt);
        else dv_dgeqrf_v03_levelx(p, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt, A, lda, T, ldt);
        dv_dlarft_connect_w03(m, vb, j, i, mt, A, lda, T, ldt);
        k += vb;
        j += vb;
    }
    return 0;
}

/* panel Q from the stored (V, T): lila-c/lila_dgeqr2_q03_3.c:3-32 */
int dv_dgeqr2_q03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                    int ldq) {
    panel_q(m - i, n, A + i + (size_t)i * lda, lda, T + i % mt + (size_t)i * ldt, ldt,
            Q + i + (size_t)i * ldq, ldq);
    return 0;
}

/* panel Q through dorgqr with tau = the T block's diagonal (its off-diagonal
 * part is not read): lila-c/lila_dgeqr2_q03_l.c:3-25 */
int dv_dgeqr2_q03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                    int ldq) {
    size_t mark = ld_ws_mark();
    double *tau = wsd((size_t)n);
    ld_T2tau(n, i, mt, T, ldt, tau);
    org_tau(m - i, n, A + i + (size_t)i * lda, lda, tau, Q + i + (size_t)i * ldq, ldq);
    ld_ws_release(mark);
    return 0;
}

/* lila-c/lila_dgeqr2_q03.c:3-19 */
int dv_dgeqr2_q03(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                  double *Q, int ldq) {
    if (p[2] == 0) return dv_dgeqr2_q03_l(m, n, i, mt, A, lda, T, ldt, Q, ldq);
    if (p[2] == 1) return dv_dgeqr2_q03_3(m, n, i, mt, A, lda, T, ldt, Q, ldq);
    return 0;
}

/* lila-c/lila_dgeqrf_q03_mt_hh.c:3-32 */
int dv_dgeqrf_q03_mt_hh(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                        int ldt, double *Q, int ldq) {
    int vb = mt - i % mt;
    if (vb > n) vb = n;
    dv_dgeqr2_q03(p, m, vb, i, mt, A, lda, T, ldt, Q, ldq);
    int j = i + vb, l = vb;
    vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
    while (vb != 0) {
        dv_dgeqr2_q03(p, m, vb, j, mt, A, lda, T, ldt, Q, ldq);
        dv_dormqrbz_w03(m, vb, l, i, j, mt, A, lda, Q, ldq, T, ldt);
        j += vb;
        l += vb;
        vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
    }
    return 0;
}

/* lila-c/lila_dgeqrf_q03_mt_l.c:3-23: dT2tau over the n columns, dorgqr */
int dv_dgeqrf_q03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                       int ldq) {
    size_t mark = ld_ws_mark();
    double *tau = wsd((size_t)n);
    ld_T2tau(n, i, mt, T, ldt, tau);
    org_tau(m - i, n, A + i + (size_t)i * lda, lda, tau, Q + i + (size_t)i * ldq, ldq);
    ld_ws_release(mark);
    return 0;
}

/* lila-c/lila_dgeqrf_q03_mt.c:3-19 */
int dv_dgeqrf_q03_mt(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                     double *Q, int ldq) {
    if (p[1] == 0) return dv_dgeqrf_q03_mt_l(m, n, i, mt, A, lda, T, ldt, Q, ldq);
    if (p[1] == 1) return dv_dgeqrf_q03_mt_hh(p, m, n, i, mt, A, lda, T, ldt, Q, ldq);
    return 0;
}

/* lila-c/lila_dgeqrf_q03_recursive.c:3-25 (leaf when n < nx) */
int dv_dgeqrf_q03_recursive(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                            int ldt, double *Q, int ldq) {
    if (n < p[3]) return dv_dgeqrf_q03_mt(p, m, n, i, mt, A, lda, T, ldt, Q, ldq);
    int n1 = n / 2, n2 = n - n1;
    dv_dgeqrf_q03_recursive(p, m, n1, i, mt, A, lda, T, ldt, Q, ldq);
    dv_dgeqrf_q03_recursive(p, m, n2, i + n1, mt, A, lda, T, ldt, Q, ldq);
    dv_dormqrbz_w03(m, n2, n1, i, i + n1, mt, A, lda, Q, ldq, T, ldt);
    return 0;
}

/* lila-c/lila_dgeqrf_q03_levelx.c:3-48 */
int dv_dgeqrf_q03_levelx(const int *p, int n_lvl, int i_lvl, const int *nb_lvl, int m, int n, int i,
                         int mt, double *A, int lda, double *T, int ldt, double *Q, int ldq) {
    int nb = nb_lvl[i_lvl];
    int vb = nb > n ? n : nb;
    int k = 0, j = i;
    if (i_lvl == n_lvl - 1) dv_dgeqrf_q03_mt(p, m, vb, i, mt, A, lda, T, ldt, Q, ldq);
    else dv_dgeqrf_q03_levelx(p, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt, A, lda, T, ldt, Q, ldq);
    k += vb;
    j += vb;
    while (j < i + n) {
        vb = (j + nb > i + n) ? (i + n - j) : nb;
        if (i_lvl == n_lvl - 1) dv_dgeqrf_q03_mt(p, m, vb, j, mt, A, lda, T, ldt, Q, ldq);
        else dv_dgeqrf_q03_levelx(p, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt, A, lda, T, ldt, Q, ldq);
        dv_dormqrbz_w03(m, vb, k, i, j, mt, A, lda, Q, ldq, T, ldt);
        k += vb;
        j += vb;
    }
    return 0;
}

/* lila-c/lila_dgeqrf_w03_appendcols.c:3-14 */
int dv_dgeqrf_w03_appendcols(const int *p, int m, int k, int n, int mt, double *A, int lda, double *Q,
                             int ldq, double *T, int ldt) {
    if (k != 0) dv_dormqrf_w03(m, n, k, 0, k, mt, A, lda, T, ldt);
    dv_dgeqrf_w03_recursive(p, m, n, k, mt, A, lda, T, ldt, Q, ldq);
    if (k != 0) dv_dlarft_connect_w03(m, n, k, 0, mt, A, lda, T, ldt);
    if (k != 0) dv_dormqrbz_w03(m, n, k, 0, k, mt, A, lda, Q, ldq, T, ldt);
    return 0;
}

/* lila-c/lila_dgeqrf_LAPACK_appendcols.c:3-30, tau in T (k + n entries);
 * A's new columns use lda (the reference passes ldq, :20) */
int dv_dgeqrf_LAPACK_appendcols(int m, int k, int n, double *A, int lda, double *Q, int ldq,
                                double *tau) {
    int ml = m - k;
    if (k != 0) orm_tau('T', m, n, k, A, lda, tau, A + (size_t)k * lda, lda);
    geqrf_tau(ml, n, A + k + (size_t)k * lda, lda, tau + k);
    org_tau(ml, n, A + k + (size_t)k * lda, lda, tau + k, Q + k + (size_t)k * ldq, ldq);
    if (k != 0) {
        ld_laset('A', k, n, 0.0, 0.0, Q + (size_t)k * ldq, ldq);
        orm_tau('N', m, n, k, A, lda, tau, Q + (size_t)k * ldq, ldq);
    }
    return 0;
}

/* lila-c/lila_dgeqrf_ker_recursive.c:3-44, tau of the columns from i at
 * T + i; A's right half uses lda (the reference passes ldq, :31) */
int dv_dgeqrf_ker_recursive(const int *p, int m, int n, int i, double *A, int lda, double *T, double *Q,
                            int ldq) {
    double *tau = T + i;
    double *Aii = A + i + (size_t)i * lda, *Qii = Q + i + (size_t)i * ldq;
    int ml = m - i;
    if (n <= p[3]) {
        geqrf_tau(ml, n, Aii, lda, tau);
        org_tau(ml, n, Aii, lda, tau, Qii, ldq);
        return 0;
    }
    int n1 = n / 2, n2 = n - n1;
    dv_dgeqrf_ker_recursive(p, m, n1, i, A, lda, T, Q, ldq);
    orm_tau('T', ml, n2, n1, Aii, lda, tau, Aii + (size_t)n1 * lda, lda);
    dv_dgeqrf_ker_recursive(p, m, n2, i + n1, A, lda, T, Q, ldq);
    ld_laset('A', n1, n2, 0.0, 0.0, Qii + (size_t)n1 * ldq, ldq);
    orm_tau('N', ml, n2, n1, Aii, lda, tau, Qii + (size_t)n1 * ldq, ldq);
    return 0;
}

/* tau_p = 2 / (1 + ||v_p(1:)||^2), then dormqr: lila-c/lila_dormqrf_w00.c:3-29
 * (C = A(i:m, j:j+n)) and lila-c/lila_dormqrf_z00.c:3-31 (C = B(i:m, j:j+n)) */
int dv_dormqrf_tau_from_v(int m, int n, int k, int i, const double *A, int lda, double *C, int ldc) {
    if (k <= 0 || n <= 0) return 0;
    size_t mark = ld_ws_mark();
    double *tau = wsd((size_t)k);
    ld_V2tau(m, k, i, A, lda, tau);
    orm_tau('T', m - i, n, k, A + i + (size_t)i * lda, lda, tau, C, ldc);
    ld_ws_release(mark);
    return 0;
}

/* one block with the full-layout T at T(i, i): lila-c/lila_dormqrf_z02.c:3-42 */
int dv_dormqrf_z02(int m, int n, int k, int i, int j, double *A, int lda, double *T, int ldt, double *B,
                   int ldb) {
    apply_block_qt(m - i, n, k, A + i + (size_t)i * lda, lda, T + i + (size_t)i * ldt, ldt,
                   B + i + (size_t)j * ldb, ldb);
    return 0;
}

/* lila-c/lila_dgeqrf_qr2.c:3-45 (lila_param[4]: 0 VRTQ, 1 VRT, 2 Q) */
int dv_dgeqrf_qr2(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                  double *Q, int ldq) {
    if (p[4] == 0) return dv_dgeqr2_w03(p, m, n, i, mt, A, lda, T, ldt, Q, ldq);
    if (p[4] == 1) return dv_dgeqr2_v03(p, m, n, i, mt, A, lda, T, ldt);
    if (p[4] == 2) return dv_dgeqr2_q03(p, m, n, i, mt, A, lda, T, ldt, Q, ldq);
    return 0;
}

/* lila-c/lila_dgeqrf_qr2_recursive.c:3-24 */
int dv_dgeqrf_qr2_recursive(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                            int ldt, double *Q, int ldq) {
    if (n <= p[3]) return dv_dgeqrf_qr2(p, m, n, i, mt, A, lda, T, ldt, Q, ldq);
    int n1 = n / 2, n2 = n - n1;
    return dv_dgeqrf_qr2_recursive2(p, m, n, i, mt, n1, n2, A, lda, T, ldt, Q, ldq);
}

/* lila-c/lila_dgeqrf_qr2_recursive2.c:3-34 */
int dv_dgeqrf_qr2_recursive2(const int *p, int m, int n, int i, int mt, int nb1, int nb2, double *A,
                             int lda, double *T, int ldt, double *Q, int ldq) {
    (void)n;
    if (p[4] == 0 || p[4] == 1) {
        dv_dgeqrf_qr2_recursive(p, m, nb1, i, mt, A, lda, T, ldt, Q, ldq);
        dv_dormqrf_w03(m, nb2, nb1, i, i + nb1, mt, A, lda, T, ldt);
        dv_dgeqrf_qr2_recursive(p, m, nb2, i + nb1, mt, A, lda, T, ldt, Q, ldq);
        dv_dlarft_connect_w03(m, nb2, i + nb1, i, mt, A, lda, T, ldt);
        if (p[4] == 0) dv_dormqrbz_w03(m, nb2, nb1, i, i + nb1, mt, A, lda, Q, ldq, T, ldt);
    } else if (p[4] == 2) {
        dv_dgeqrf_qr2_recursive(p, m, nb1, i, mt, A, lda, T, ldt, Q, ldq);
        dv_dgeqrf_qr2_recursive(p, m, nb2, i + nb1, mt, A, lda, T, ldt, Q, ldq);
        dv_dormqrbz_w03(m, nb2, nb1, i, i + nb1, mt, A, lda, Q, ldq, T, ldt);
    }
    return 0;
}

/* ---- workspace queries: the reference's closed forms (lila-c/lila_query_*.c) */

int dq_dormqrf_w03(int m, int n, int k, int i, int mt) {
    int vb = mt - i % mt;
    if (vb > k) vb = k;
    int w = n * mt, ml = m - i, jj = 1;
    for (;;) {
        if (w < (ml - vb) * n) w = (ml - vb) * n;
        if (jj + vb - 1 >= k) break;
        ml -= vb;
        jj += vb;
        vb = (jj + mt - 1 <= k) ? mt : (k - jj + 1);
    }
    return w;
}

int dq_dormqrbz_w03(int m, int n, int k, int i, int mt) {
    int vb = (i + k) % mt;
    if (vb == 0) vb = mt;
    if (vb > k) vb = k;
    int w = n * k, ml = m - (i + k - vb), jj = 1;
    while (jj + vb - 1 < k) {
        if (w < n * vb) w = n * vb;
        if (w < (ml - vb) * k) w = (ml - vb) * k;
        jj += vb;
        vb = (jj + mt - 1 <= k) ? mt : (k - jj + 1);
        ml += vb;
    }
    return w;
}

static int dq_dgeqr2(const int *p, int n) {
    if (p[2] == 0) return n * n;
    if (p[2] == 1) return n * n;
    if (p[2] == 2) return n;
    return 0;
}

int dq_dgeqrf_w03_mt(const int *p, int m, int n, int i, int mt, int lwork) {
    int w = lwork, q;
    if (p[1] == 0) {
        if (w < n * n) w = n * n;
    } else if (p[1] == 1) {
        int vb = mt - i % mt;
        if (vb > n) vb = n;
        q = dq_dgeqr2(p, vb);
        if (w < q) w = q;
        int j = i + vb, l = vb;
        vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
        while (vb != 0) {
            q = dq_dormqrf_w03(m, vb, l, i, mt);
            if (w < q) w = q;
            q = dq_dgeqr2(p, vb);
            if (w < q) w = q;
            q = dq_dormqrbz_w03(m, vb, l, i, mt);
            if (w < q) w = q;
            j += vb;
            l += vb;
            vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
        }
    } else if (p[1] == 2) {
        int j = i, vb = mt - i % mt;
        if (vb > n) vb = n;
        while (vb != 0) {
            q = vb * (j - i) + vb;
            if (w < q) w = q;
            j += vb;
            vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
        }
    }
    return w;
}

int dq_dgeqrf_w03_recursive(const int *p, int m, int n, int i, int mt, int lwork) {
    int w = lwork, q;
    if (n <= p[3]) {
        q = dq_dgeqrf_w03_mt(p, m, n, i, mt, w);
        if (w < q) w = q;
    } else {
        int n1 = n / 2, n2 = n - n1;
        q = dq_dgeqrf_w03_recursive(p, m, n1, i, mt, w);
        if (w < q) w = q;
        q = dq_dormqrf_w03(m, n2, n1, i, mt);
        if (w < q) w = q;
        q = dq_dgeqrf_w03_recursive(p, m, n2, i + n1, mt, w);
        if (w < q) w = q;
        q = dq_dormqrbz_w03(m, n2, n1, i, mt);
        if (w < q) w = q;
    }
    return w;
}

int dq_dgeqrf_w03_levelx(const int *p, int n_lvl, int i_lvl, const int *nb_lvl, int m, int n, int i,
                         int mt) {
    int w = 0, q;
    int nb = nb_lvl[i_lvl];
    int vb = nb > n ? n : nb;
    int k = 0, j = i;
    if (i_lvl == n_lvl - 1) q = dq_dgeqrf_w03_mt(p, m, n, i, mt, w);
    else q = dq_dgeqrf_w03_levelx(p, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt);
    if (q > w) w = q;
    k += vb;
    j += vb;
    while (j < i + n) {
        vb = (j + nb > i + n) ? (i + n - j) : nb;
        q = dq_dormqrf_w03(m, vb, k, i, mt);
        if (q > w) w = q;
        if (i_lvl == n_lvl - 1) q = dq_dgeqrf_w03_mt(p, m, n, i, mt, w);
        else q = dq_dgeqrf_w03_levelx(p, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt);
        if (q > w) w = q;
        q = dq_dormqrbz_w03(m, vb, k, i, mt);
        if (q > w) w = q;
        k += vb;
        j += vb;
    }
    return w;
}
