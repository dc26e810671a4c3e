/*
 * lila_oracle.h -- CPU oracle for the LILA Householder-QR hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.  The
 * product (liblila_hip.so) never links, calls or falls back to it.
 *
 * It restates, in plain C, the algorithms of the reference's lila-c/ w03
 * snapshot (each function cites the reference file:line it follows) on top of
 * a small built-in BLAS/LAPACK subset (naive loops, LAPACK's published
 * algorithms for dlarfg / dlarft / dgeqr2 / dorg2r / dpotrf).  Every symbol
 * carries the prefix "ora_" so it can live in one process with the GPU
 * library, whose symbols keep the reference names.
 *
 * ora_use_host_blas() optionally routes the level-2/3 BLAS calls (dgemm,
 * dtrmm, dtrsm, dsyrk, dger) to a host BLAS loaded with dlopen; that mode is
 * used for the CPU-baseline timing, never for pinning.
 */
#ifndef LILA_ORACLE_H
#define LILA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- built-in BLAS/LAPACK subset (column-major, LAPACK argument order) -- */
void ora_dgemm(char ta, char tb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb,
               double beta, double *C, int ldc);
void ora_dtrmm(char side, char uplo, char trans, char diag, int m, int n,
               double alpha, const double *A, int lda, double *B, int ldb);
void ora_dtrsm(char side, char uplo, char trans, char diag, int m, int n,
               double alpha, const double *A, int lda, double *B, int ldb);
void ora_dsyrk(char uplo, char trans, int n, int k, double alpha,
               const double *A, int lda, double beta, double *C, int ldc);
void ora_dger(int m, int n, double alpha, const double *x, int incx,
              const double *y, int incy, double *A, int lda);
void ora_dlarfg(int n, double *alpha, double *x, int incx, double *tau);
void ora_dlarft_fc(int n, int k, const double *V, int ldv, const double *tau,
                   double *T, int ldt);
void ora_dgeqr2(int m, int n, double *A, int lda, double *tau);
void ora_dorg2r(int m, int n, int k, double *A, int lda, const double *tau);
int  ora_dpotrf_u(int n, double *A, int lda);
void ora_dlacpy(char uplo, int m, int n, const double *A, int lda, double *B, int ldb);
void ora_dlaset(char uplo, int m, int n, double alpha, double beta, double *A, int lda);
double ora_dlange_f(int m, int n, const double *A, int lda);

/* Route level-2/3 BLAS to a host library (path, symbol prefix such as
 * "scipy_" or "").  Returns 0 on success.  ora_use_builtin_blas() reverts. */
int  ora_use_host_blas(const char *path, const char *prefix);
void ora_use_builtin_blas(void);

/* ---- LILA w03 routines (same argument lists as lila-c/lila.h) ---------- */
int ora_dgeqr3(int m, int n, double *A, int lda, double *T, int ldt);
int ora_lila_dlarft_w03(int m, int n, int i, int mt, double *A, int lda,
                        double *T, int ldt, double *tau);
int ora_lila_dlarft_connect_w03(int m, int n, int i, int j, int mt,
                                double *A, int lda, double *T, int ldt);
int ora_lila_dormqrf_w03(int m, int n, int k, int i, int j, int mt,
                         double *A, int lda, double *T, int ldt,
                         double *work, int lwork);
int ora_lila_dormqrf_z03(int m, int n, int k, int i, int j, int mt,
                         double *A, int lda, double *T, int ldt,
                         double *B, int ldb, double *work, int lwork);
int ora_lila_dormqrbz_w03(int m, int n, int k, int i, int j, int mt,
                          double *A, int lda, double *Q, int ldq,
                          double *T, int ldt, double *work, int lwork);
int ora_lila_dgeqr2_w03(int *lila_param, int m, int n, int i, int mt,
                        double *A, int lda, double *T, int ldt,
                        double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqr2_w03_3(int m, int n, int i, int mt, double *A, int lda,
                          double *T, int ldt, double *Q, int ldq,
                          double *work, int lwork);
int ora_lila_dgeqr2_w03_l(int m, int n, int i, int mt, double *A, int lda,
                          double *T, int ldt, double *Q, int ldq,
                          double *work, int lwork);
int ora_lila_dgeqr2_w03_hr(int m, int n, int i, int mt, double *A, int lda,
                           double *T, int ldt, double *Q, int ldq,
                           double *work, int lwork);
int ora_lila_dgeqrf_w03_mt(int *lila_param, int m, int n, int i, int mt,
                           double *A, int lda, double *T, int ldt,
                           double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqrf_w03_mt_hh(int *lila_param, int m, int n, int i, int mt,
                              double *A, int lda, double *T, int ldt,
                              double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqrf_w03_mt_l(int m, int n, int i, int mt, double *A, int lda,
                             double *T, int ldt, double *Q, int ldq,
                             double *work, int lwork);
int ora_lila_dgeqrf_w03_mt_hr(int m, int n, int i, int mt, double *A, int lda,
                              double *T, int ldt, double *Q, int ldq,
                              double *work, int lwork);
int ora_lila_dgeqrf_w03_recursive(int *lila_param, int m, int n, int i, int mt,
                                  double *A, int lda, double *T, int ldt,
                                  double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqrf_w03_levelx(int *lila_param, int n_lvl, int i_lvl,
                               int *nb_lvl, int m, int n, int i, int mt,
                               double *A, int lda, double *T, int ldt,
                               double *Q, int ldq, double *work, int lwork);
int ora_lila_ormhr2_w03_hr(int m, int n, int i, int j, int l, int mt,
                           double *A, int lda, double *T, int ldt,
                           double *Q, int ldq, double *work, int lwork,
                           double *S);
int ora_lila_dorgh2_3(int m, int n, double *A, int lda, double *T, int ldt,
                      double *Q, int ldq, double *work, int lwork, double *S);
int ora_lila_dorgh2(int m, int n, int i, int j, int mt, double *A, int lda,
                    double *T, int ldt, double *Q, int ldq, double *work,
                    int lwork, double *S);
int ora_lila_dorghr(int m, int n, int i, int mt, double *A, int lda,
                    double *T, int ldt, double *Q, int ldq, double *work,
                    int lwork, int *S);
int ora_lila_dgetrf_b(int m, int n, int i, int mt, double *A, int lda,
                      double *Q, int ldq, double *work, int lwork);
int ora_lila_dT2tau_w03(int m, int n, int i, int mt, double *T, int ldt,
                        double *tau);
int ora_lila_dV2tau_w03(int m, int n, int i, int mt, double *A, int lda,
                        double *tau);

/* VRT-only (v03), Q from (V,T) (q03), append-columns, LAPACK-style (ker,
 * w00/z00) and qr2 families (same argument lists as lila-c/lila.h) */
int ora_lila_dgeqr2_v03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                          double *work, int lwork);
int ora_lila_dgeqr2_v03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                          double *work, int lwork);
int ora_lila_dgeqr2_v03(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                        double *T, int ldt, double *work, int lwork);
int ora_lila_dgeqrf_v03_mt_hh(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                              double *T, int ldt, double *work, int lwork);
int ora_lila_dgeqrf_v03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                             double *work, int lwork);
int ora_lila_dgeqrf_v03_mt(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                           double *T, int ldt, double *work, int lwork);
int ora_lila_dgeqrf_v03_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                  double *T, int ldt, double *work, int lwork);
int ora_lila_dgeqrf_v03_levelx(int *lila_param, int n_lvl, int i_lvl, int *nb_lvl, int m, int n,
                               int i, int mt, double *A, int lda, double *T, int ldt, double *work,
                               int lwork);
int ora_lila_dgeqr2_q03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                          double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqr2_q03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                          double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqr2_q03(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                        double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqrf_q03_mt_hh(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                              double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqrf_q03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                             double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqrf_q03_mt(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                           double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqrf_q03_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                  double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqrf_q03_levelx(int *lila_param, int n_lvl, int i_lvl, int *nb_lvl, int m, int n,
                               int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                               int ldq, double *work, int lwork);
int ora_lila_dgeqrf_w03_appendcols(int *lila_param, int m, int k, int n, int mt, double *A, int lda,
                                   double *Q, int ldq, double *T, int ldt, double *work, int lwork);
int ora_lila_dgeqrf_LAPACK_appendcols(int m, int k, int n, int mt, double *A, int lda, double *Q,
                                      int ldq, double *T, int ldt, double *work, int lwork);
int ora_lila_dgeqrf_ker_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                  double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int ora_lila_dormqrf_w00(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T,
                         int ldt, double *work, int lwork);
int ora_lila_dormqrf_z00(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *B,
                         int ldb, double *T, int ldt, double *work, int lwork);
int ora_lila_dormqrf_z02(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T,
                         int ldt, double *B, int ldb, double *work, int lwork);
int ora_lila_dgeqrf_qr2(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                        int ldt, double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqrf_qr2_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                  double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int ora_lila_dgeqrf_qr2_recursive2(int *lila_param, int m, int n, int i, int mt, int nb1, int nb2,
                                   double *A, int lda, double *T, int ldt, double *Q, int ldq,
                                   double *work, int lwork);

/* workspace queries (return doubles needed; lwork is the running maximum) */
int ora_lila_query_dormqrf_w03(int m, int n, int k, int i, int j, int mt,
                               double *A, int lda, double *T, int ldt,
                               double *work, int lwork);
int ora_lila_query_dormqrbz_w03(int m, int n, int k, int i, int j, int mt,
                                double *A, int lda, double *Q, int ldq,
                                double *T, int ldt, double *work, int lwork);
int ora_lila_query_dgeqrf_w03_mt(int *lila_param, int m, int n, int i, int mt,
                                 double *A, int lda, double *T, int ldt,
                                 double *Q, int ldq, double *work, int lwork);
int ora_lila_query_dgeqrf_w03_recursive(int *lila_param, int m, int n, int i,
                                        int mt, double *A, int lda, double *T,
                                        int ldt, double *Q, int ldq,
                                        double *work, int lwork);
int ora_lila_query_dgeqrf_w03_levelx(int *lila_param, int n_lvl, int i_lvl,
                                     int *nb_lvl, int m, int n, int i, int mt,
                                     double *A, int lda, double *T, int ldt,
                                     double *Q, int ldq, double *work,
                                     int lwork);

/* least squares through dgeqr3 + (V,T) + R solve, per
 * least-squares/main_ls_qr3.c:94-119.  B (m x k) is overwritten; its first
 * n rows hold X on return.  work: n*k doubles, T: n*n. */
int ora_ls_qr3(int m, int n, int k, double *A, int lda, double *T, int ldt,
               double *B, int ldb, double *work);

/* LAPACK dgels('N'), m >= n, as dgeqr2 + dorm2r + back substitution: the
 * per-problem host baseline of the batched least-squares workload
 * (least-squares/main_lapack_ref.c:100).  tau: n doubles out. */
int ora_dgels_hh(int m, int n, int k, double *A, int lda, double *B, int ldb, double *tau);

/* the recursive QR library of qr3/src (qr2.h): qr2_dgeqr3_R.c, qr2_dlarft3.c,
 * qr2_dorgqr3.c restated */
int ora_qr2_dgeqr3_R(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr);
int ora_qr2_dlarft3(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int ora_qr2_dorgqr3(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau);
/* the UT / ISW / V2T / VT2Q / VS2Q forms and the aux routines of qr3/src */
int ora_qr2_dgeqr3_R_UT(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr);
int ora_qr2_dgeqr3_R_ISW(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr);
int ora_qr2_dgeqr3_R_UT_ISW(int m, int n, double *A, int lda, double *T, int ldt, double *R,
                            int ldr);
int ora_qr2_dlarft3_UT(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int ora_qr2_dlarft3_ISW(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int ora_qr2_dlarft3_UT_ISW(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int ora_qr2_dlarft3_ISW_V2T(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int ora_qr2_dlarft3_ISW_V2T_UT(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int ora_qr2_dorgqr3_UT(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau);
int ora_qr2_dorgqr3_VT2Q(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau);
int ora_qr2_dorgqr3_VS2Q(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau);
int ora_qr2_aux_dV2N(int n, double *T, int ldt);
int ora_qr2_aux_dN2T(int n, double *tau, double *T, int ldt);
int ora_qr2_aux_dVS2Q(int m, int n, double *Q, int ldq);
int ora_qr2_aux_dVT2Q(int m, int n, double *Q, int ldq);
int ora_qr2_aux_dmLUinA(int n, double *A, int lda);
int ora_qr2_aux_dV2tau(int m, int n, double *A, int lda, double *tau);
int ora_qr2_aux_dULTinU(int n, double *L, int ldl, double *U, int ldu);
int ora_qr2_aux_dApUBTinA(int m, int n, double *A, int lda, double *U, int ldu, double *B, int ldb);
int ora_qr2_aux_dUinvLTinU(int n, double *L, int ldl, double *U, int ldu);
int ora_qr2_aux_dlarf_wrapper(char side, int m, int n, double *V, int incv, double tau, double *C,
                              int ldc, double *work);

/* ---- checkers (lila-c-lila_param/lila/src/lila_test_*.c) -------------- */
double ora_test_qq_orth_1(int m, int n, int i, const double *Q, int ldq);
double ora_test_qq_orth_full(int m, int n, int i, const double *Q, int ldq);
double ora_test_qr_repres_1(int m, int n, int i, const double *A, int lda,
                            const double *Q, int ldq, const double *R, int ldr);
double ora_test_r_repres_2(int m, int n, int k, int i, int mt,
                           const double *As, int lda, const double *R, int ldr,
                           const double *T, int ldt);

/* ---- input generators shared with the GPU tests ------------------------ */
/* splitmix64 stream -> Uniform(lo,hi) or N(0,1) (Box-Muller), column-major
 * m x n with leading dimension ld. */
void ora_fill_uniform(unsigned long long seed, int m, int n, double *A, int ld,
                      double lo, double hi);
void ora_fill_normal(unsigned long long seed, int m, int n, double *A, int ld);

#ifdef __cplusplus
}
#endif
#endif
