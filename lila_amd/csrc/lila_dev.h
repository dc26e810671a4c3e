/*
 * lila_dev.h -- internal C interface between the host drivers (plain C,
 * drivers.c / api.c) and the HIP side (runtime.hip and the kernel files).
 *
 * Every ld_* routine enqueues work on the library stream and returns
 * immediately; all pointers are device pointers into column-major FP64
 * matrices.  Nothing here allocates per call except through the workspace
 * arena (ld_ws_*), which the API layer sizes before a call tree runs.
 */
#ifndef LILA_DEV_H
#define LILA_DEV_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* structure of a stored operand (applied to stored row r, column c before
 * any transpose): general, unit lower trapezoidal (Householder V: 0 above,
 * 1 on the diagonal), or upper triangular (T, R: 0 below the diagonal). */
/* LD_UPPER_Z: upper triangular with exact zeros stored below the diagonal
 * (read as a general matrix; the zeros only shorten the K range) */
enum { LD_GEN = 0, LD_VUNIT = 1, LD_UPPER = 2, LD_LOWER = 3, LD_UUNIT = 4, LD_UPPER_Z = 5 };

typedef struct {
    int m, n, k;
    double alpha, beta;
    const double *A; int lda; int ta; int sa;
    const double *B; int ldb; int tb; int sb;
    double *C; int ldc;
    int zero_rows;  /* rows [0, zero_rows) of C take beta = 0 */
    int c_upper;    /* 1: only C(i,j) with i <= j is written (SYRK form) */
} ld_gemm_t;

/* C = alpha op(A) op(B) + beta C on FP64 MFMA (split-K chosen inside) */
void ld_gemm(const ld_gemm_t *g);
/* the same emulated on int8 MFMA by the CRT scheme with kb-bit operands
 * (gemm_emu.hip): general op(A) = A or A^T, B not transposed, K <= 65536;
 * 0 when done, -1 when the product is not one it takes */
int ld_gemm_emu(const ld_gemm_t *g, int kb);
int ld_emu_moduli(int K, int kb);
static inline void ld_gemm_simple(int ta, int tb, int m, int n, int k, double alpha,
                                  const double *A, int lda, int sa,
                                  const double *B, int ldb, int sb,
                                  double beta, double *C, int ldc) {
    ld_gemm_t g;
    g.m = m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
    g.A = A; g.lda = lda; g.ta = ta; g.sa = sa;
    g.B = B; g.ldb = ldb; g.tb = tb; g.sb = sb;
    g.C = C; g.ldc = ldc; g.zero_rows = 0; g.c_upper = 0;
    ld_gemm(&g);
}

/* W (m x n, ldw) = op(T) (alpha op(A) op(B)), op(T) = T^T (transT) or T
 * with T upper m x m, when the product is split over K and m <= 64: the
 * split partials and one launch that sums and multiplies; returns 1 and
 * does nothing otherwise (beta, C, zero_rows, c_upper of g unused) */
int ld_gemm_tmul(const ld_gemm_t *g, const double *T, int ldt, int transT, double *W, int ldw);
/* element-wise helpers */
void ld_lacpy(int uplo, int m, int n, const double *A, int lda, double *B, int ldb);
void ld_laset(int uplo, int m, int n, double offdiag, double diag, double *A, int lda);
void ld_add_diag(int n, double *A, int lda, double v);
/* A(i,i) = v[i*incv], or 1 / v[i*incv] when recip */
void ld_set_diag(int n, const double *v, int incv, int recip, double *A, int lda);
void ld_sub(int m, int n, const double *W, int ldw, double *C, int ldc); /* C -= W */

/* Fused Householder panel for n <= LD_PANEL_MAX columns (ml >= n):
 * V and R in A, the n x n upper T at T (ldt), and when Q != NULL the
 * explicit panel Q = [I;0] - V T V1^T (ml x n) at Q (ldq). */
#define LD_PANEL_MAX 32
void ld_panel_qr(int ml, int n, double *A, int lda, double *T, int ldt,
                 double *Q, int ldq);
/* the 64-wide Cholesky-QR2 panel (32 < n <= 64, no explicit Q, panel mode
 * 1): 0 launched, -1 not taken (nothing launched) */
int ld_panel_qr64(int ml, int n, double *A, int lda, double *T, int ldt);
/* the two block steps of a dgeqr3 node whose halves are at most 64 wide,
 * one launch each (node_qt.hip); 0 launched, -1 not taken (nothing
 * launched).  update: C (ml x n2) -= V (T1^T (V^T C)), V ml x n1 unit lower;
 * join: T12 (n1 x n2) = -T1 (V1^T V2) T2, V1 ml2 x n1, V2 ml2 x n2 unit lower */
int ld_node_update64(int ml, int n1, int n2, const double *V, int ldv, const double *T1, int ldt,
                     double *C, int ldc);
int ld_node_join64(int ml2, int n1, int n2, const double *V1, int ldv1, const double *V2, int ldv2,
                   const double *T1, int ldt1, const double *T2, int ldt2, double *T12, int ldt12);
/* tall panels (taller than the workgroups' LDS slabs hold): */
int ld_panel_is_tall(int ml, int n);
/* the streamed Cholesky-QR2 panel without its column-loop fallback: 0 done,
 * 1 declined (acceptance test failed, panel untouched); waits for it */
int ld_panel_qr_tall_cqr(int ml, int n, double *A, int lda, double *T, int ldt, double *Q, int ldq);
/* TSQR: slab count of an M x nb level, the leaf QRs of a level (V, R in
 * place, T_s at Ts + s*32*32 (ld 32), stacked R's into Y (S nb x nb, ld ldy)),
 * and the top-down expansion of a level's explicit Q rows (Xq: the level
 * above's Q, NULL for the identity) */
int ld_tsqr_slabs(int M, int nb);
void ld_tsqr_leaf(int M, int nb, double *X, int ldx, int S, double *Ts, double *Y, int ldy);
void ld_tsqr_expand(int M, int nb, int S, const double *V, int ldv, const double *Ts,
                    const double *Xq, int ldx, double *Qo, int ldq);
/* the same, always column-by-column (dgeqr3's own arithmetic) */
void ld_panel_qr_exact(int ml, int n, double *A, int lda, double *T, int ldt,
                       double *Q, int ldq);

/* dlarft ('F','C') for k <= LD_PANEL_MAX... any k: T = fwd column T from V
 * (ml x k, unit lower) and tau[k] (device) */
void ld_larft(int ml, int k, const double *V, int ldv, const double *tau,
              double *T, int ldt);
/* T (n x n upper) from the strict upper part of N = V^T V held in T and
 * tau: the dlarft recurrence, T12 = -T11 N12 T22 between halves */
void ld_n2t(int n, const double *tau, double *T, int ldt);
/* copy diag of T blocks into tau (n entries, global start i) */
void ld_T2tau(int n, int i, int mt, const double *T, int ldt, double *tau);
/* tau_c = 2 / (1 + ||A(g+1:m, g)||^2), g = i + c */
void ld_V2tau(int m, int n, int i, const double *A, int lda, double *tau);

/* Triangular solve with op(A) (side 'L'/'R', uplo 'U'/'L', trans 'N'/'T',
 * diag 'N'/'U'), in place on B (m x n), alpha scaling. */
void ld_trsm(char side, char uplo, char trans, char diag, int m, int n,
             double alpha, const double *A, int lda, double *B, int ldb);
/* Triangular multiply B = alpha op(A) B or alpha B op(A), in place. */
void ld_trmm(char side, char uplo, char trans, char diag, int m, int n,
             double alpha, const double *A, int lda, double *B, int ldb);
/* G(upper) = A^T A for n <= 256 with 3-term split-BF16 MFMA products and
 * FP32 accumulation (gram.hip); -1 when n > 256 */
int ld_gram_bf16x3(int m, int n, const double *A, int lda, double *G, int ldg);
/* the same in FP64 MFMA arithmetic */
int ld_gram_f64(int m, int n, const double *A, int lda, double *G, int ldg);
/* C(upper) = alpha A^T A + beta C(upper)  (A is k x n) */
void ld_syrk_ut(int n, int k, double alpha, const double *A, int lda,
                double beta, double *C, int ldc);
/* upper Cholesky in place; returns through the info slot (ld_info) */
void ld_potrf_u(int n, double *A, int lda);
double ld_orth_defect(int n, const double *G, int ldg); /* ||I - G||_F, syncs */
double ld_sym_defect_norm2_est(int n, const double *G, int ldg, int iters); /* <= ||I - G||_2, syncs */

/* sign-chosen unpivoted LU of the n x n block Tq in place (Householder
 * reconstruction): S[k] = -1 if |1 - t_kk| < |1 + t_kk| (column negated),
 * then t_kk -= 1, column scaled, rank-1 update. S is a device double array. */
void ld_lu_sign(int n, double *Tq, int ldt, double *S);
/* negate rows r in [0, nrows) of the n columns... helpers driven by S:
 *   ld_flip_cols: column c of X (m rows) negated when S[c] == -1
 *   ld_flip_rows_upper: row k of X (cols [k, n)) negated when S[k] == -1
 *   ld_flip_rows: row k of X (all n cols) negated when S[k] == -1 */
void ld_flip_cols(int m, int n, double *X, int ldx, const double *S);
void ld_flip_rows_upper(int n, double *X, int ldx, const double *S);
void ld_flip_rows(int nrows, int n, double *X, int ldx, const double *S);
/* move strict lower of Tq into A's strict lower and zero it in Tq */
void ld_move_lower(int n, double *Tq, int ldt, double *A, int lda);
/* dgetrf_b body (c3/lila_dgetrf_b.c): LU of [A1 D + I; A2 D] style */
void ld_getrf_b(int ml, int n, double *A, int lda, double *Q, int ldq, double *U);

/* generators (same splitmix64 stream as the oracle's) */
void ld_fill_uniform(unsigned long long seed, int m, int n, double *A, int ld,
                     double lo, double hi);
void ld_fill_normal(unsigned long long seed, int m, int n, double *A, int ld);


/* ---- runtime ------------------------------------------------------------ */
int  ld_init(void);               /* 0 on success */
void *ld_stream(void);            /* hipStream_t of the current lane */
void ld_set_stream(void *s);      /* lane 0's stream (the library stream) */
/* lanes: lane 0 runs on the library stream, lane 1 on an internal stream
 * with its own workspace arena.  ld_lane_fork makes lane 1 wait for the
 * work enqueued on lane 0 so far, ld_lane_join the reverse; ld_set_lane
 * selects where the next ld_* calls go.  ld_lanes() == 1 keeps everything
 * on lane 0. */
int  ld_lanes(void);
int  ld_set_lanes(int n);         /* returns the previous count */
void ld_set_lane(int lane);
void ld_lane_fork(void);
void ld_lane_join(void);
int  ld_is_device_ptr(const void *p);
void ld_ptr_diag(const void *p, int *err, int *type);
void *ld_ws_alloc(size_t bytes);  /* bump allocation from the arena; NULL (and
                                     ld_failed() true) when no chunk can be had */
int  ld_stream_capturing(void);  /* the library stream is being captured */
int  ld_failed(void);             /* an allocation of the current call failed:
                                     launchers and copies return at once */
void ld_clear_failed(void);
size_t ld_ws_mark(void);
void ld_ws_release(size_t mark);
int  ld_ws_reserve(size_t bytes); /* ensure one free chunk of this size */
void ld_ws_consolidate(void);     /* fold chunks into one (call when idle) */
void *ld_dev_malloc(size_t bytes);
void ld_dev_free(void *p);
void ld_memcpy_h2d(void *dst, const void *src, size_t bytes);
void ld_memcpy_d2h(void *dst, const void *src, size_t bytes);
int  ld_sync(void);
/* status: 0 ok; >0 numerical info (e.g. potrf column); <0 errors */
void ld_set_info(int info);
int  ld_get_info(void);
void ld_clear_info(void);
/* device-side error flags (barrier timeouts, potrf breakdown), read back
 * and folded into the info slot at ld_sync */
int  ld_check_device_flags(void);
void ld_fatal(const char *msg);

/* opt-in launch profiling (HIP events on the library stream) */
int  ld_prof_begin(const char *kernel_class);  /* -1 when profiling is off */
void ld_prof_end(int handle, double flops, double bytes);

/* tunables (autotuner writes these) */
typedef struct {
    int panel_rows_per_wg;   /* target rows per workgroup, column-by-column panel */
    int panel_rows_cqr;      /* target rows per workgroup, Cholesky-QR2 panel */
    int gemm_min_wgs;        /* split-K target occupancy */
    int trsm_nb;             /* diagonal block of trsm / potrf */
    int gemm_bm, gemm_bn;    /* default (large) GEMM tile */
    int gemm_small_tiles;    /* below this many large tiles use 64 x 64 */
    int panel_mode;          /* 0 column Householder, 1 Cholesky-QR2 + reconstruction */
    int gemm_lowp;           /* 1: GEMMs on FP32 MFMA (operands rounded to FP32) */
    int gemm_bm_t, gemm_bn_t; /* tile of products with op(A) = A^T (W = V^T C) */
    int gram_prec;           /* Cholesky-QR of the hr paths: 0 one FP64 pass; 1, 2, 3 two
                                passes whose first Gram is FP64, FP32 MFMA, split BF16 */
    int gemm_small_tiles_t;  /* gemm_small_tiles for products with op(A) = A^T */
    int batch_grid;          /* batched least squares: workgroups of the persistent launch
                                (0: one per CU at the kernel's occupancy) */
    int panel_nb;            /* panel width: dgeqr3 recursion cutoff (8..32, or 64: widths
                                33..64 to the 64-wide Cholesky-QR2 panel where it fits;
                                0 = 32) */
    int tune_form;           /* lila_autotune times the plain (0) or mixed (1) form */
    int gemm_kmin;           /* split-K: rows of K per split at least (0 = 64) */
    int gemm_v2;             /* 1: general products on the 128 x 128 direct-to-LDS kernel */
    int q_after;             /* VRTQ recursion: 1 builds Q after the VRT recursion by whole T
                                blocks, 0 at every node as the reference's call tree does */
    int node_fused;          /* dgeqr3 nodes with halves <= 64 wide: update and T join in
                                one launch each (node_qt.hip) */
    int gemm_rocblas;        /* 1: plain NN interiors of the large products through rocBLAS,
                                0 (default): on gemm2_f64_kernel */
} ld_tune_t;
ld_tune_t *ld_tune(void);

/* batched least squares (batched.hip): one persistent launch over many
 * independent problems; the plan holds the problem table in HBM */
typedef struct ld_batch_plan ld_batch_plan;
ld_batch_plan *ld_batch_plan_create(int nprob, const int *m, const int *n, const int *k,
                                    double *const *A, const int *lda, double *const *B,
                                    const int *ldb);
int ld_batch_plan_run(ld_batch_plan *pl, int use_graph);
double ld_batch_plan_flops(const ld_batch_plan *pl);
void ld_batch_plan_destroy(ld_batch_plan *pl);
/* the same launch with the problem table, queue order and counter in a
 * static device array instead of allocated memory (lila_dgels_batched_noalloc) */
int ld_batch_noalloc_setup(void);
size_t ld_batch_noalloc_max(void);
int ld_batch_noalloc_run(int nprob, const int *m, const int *n, const int *k, double *const *A,
                         const int *lda, double *const *B, const int *ldb);

#ifdef __cplusplus
}
#endif
#endif
