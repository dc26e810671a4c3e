/*
 * lila.h -- drop-in C ABI of the LILA Householder-QR hot path, implemented
 * on one AMD MI355X (gfx950) by liblila_hip.so.
 *
 * The prototypes below keep the argument lists of the reference header
 * (lila-c/lila.h, plus lila_dorghr / lila_dgetrf_b from c3/lila.h:13,43), so
 * a caller of the reference links against liblila_hip.so unchanged.  Each
 * entry cites the reference declaration it replaces.
 *
 * Conventions (as in the reference): column-major FP64; global index i means
 * the routine works on A(i:m-1, i:i+n-1); T is mt x (i+n) with ldt = mt and
 * mt-aligned upper-triangular diagonal blocks (block b holds reflectors
 * [b*mt, (b+1)*mt)); Q is m x (i+n); work / lwork keep their meaning but the
 * library uses its own device workspace, so any work pointer (or NULL) is
 * accepted.
 *
 * Pointers may be host memory (the library copies the touched span in and
 * out and returns after the copy-out) or device memory (hipMalloc / a torch
 * tensor: the call is asynchronous on the library stream, like a BLAS call).
 *
 * Return value: 0 on success, a positive LAPACK-style info on numerical
 * breakdown (Cholesky not positive definite: the 1-based column), negative
 * on a device or argument error: -1 no device / bad arguments, -2 device
 * memory for staging or workspace could not be allocated (nothing was
 * computed), -4 a panel's grid barrier timed out, -6 the library stream is
 * being captured into a hipGraph (the dense entry points refuse capture;
 * the batched plan replays a graph of its own).  The reference always
 * returns 0.
 * Not thread-safe: one calling thread per process.
 */
#ifndef LILA_H_MI355X
#define LILA_H_MI355X

#ifdef __cplusplus
extern "C" {
#endif

/* lila-c/lila.h:9 -- recursive panel QR (Elmroth-Gustavson); T n x n upper */
int dgeqr3(int m, int n, double *A, int lda, double *T, int ldt);

/* lila-c/lila.h:30 -- A(:, j:j+n) <- Q^T A(:, j:j+n), k reflectors from i */
int lila_dormqrf_w03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T,
                     int ldt, double *work, int lwork);
/* lila-c/lila.h:31 -- B(:, j:j+n) <- Q^T B(:, j:j+n) */
int lila_dormqrf_z03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T,
                     int ldt, double *B, int ldb, double *work, int lwork);
/* lila-c/lila.h:34 -- Q(:, j:j+n) <- H_i..H_{i+k-1} [0; Q(:, j:j+n)].
 * Positional order (A, Q, T) is that of the definition
 * (lila-c/lila_dormqrbz_w03.c:3); the reference header swaps the names. */
int lila_dormqrbz_w03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *Q,
                      int ldq, double *T, int ldt, double *work, int lwork);

/* lila-c/lila.h:38 -- sign-chosen LU reconstruction of block [j, j+n) */
int lila_dorgh2(int m, int n, int i, int j, int mt, double *A, int lda, double *T, int ldt,
                double *Q, int ldq, double *work, int lwork, double *S);
/* lila-c/lila.h:39 */
int lila_dorgh2_3(int m, int n, double *A, int lda, double *T, int ldt, double *Q, int ldq,
                  double *work, int lwork, double *S);
/* lila-c/lila.h:40 */
int lila_ormhr2_w03_hr(int m, int n, int i, int j, int l, int mt, double *A, int lda, double *T,
                       int ldt, double *Q, int ldq, double *work, int lwork, double *S);
/* c3/lila.h:43 -- blocked Householder reconstruction of A(i:m, i:i+n) */
int lila_dorghr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                int ldq, double *work, int lwork, int *S);
/* c3/lila.h:13 -- sign-chosen unpivoted LU; work holds R (n x n, ld n) */
int lila_dgetrf_b(int m, int n, int i, int mt, double *A, int lda, double *Q, int ldq, double *work,
                  int lwork);

/* lila-c/lila.h:43 */
int lila_dlarft_connect_w03(int m, int n, int i, int j, int mt, double *A, int lda, double *T,
                            int ldt);
/* lila-c/lila.h:44 */
int lila_dlarft_w03(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                    double *tau);

/* lila-c/lila.h:46-56.  lila_param[0] (mode, unused by the reference's
 * drivers) selects, in lila_dgeqrf_w03_recursive, 0: the FP64 recursion;
 * 1: mixed precision -- FP32-MFMA Householder R0 as a preconditioner, FP64
 * Cholesky-QR on A R0^-1 (a second pass unless the first Gram G shows
 * ||I - G||_2 <= 0.5), Householder reconstruction to the same V, R, T, Q
 * outputs (lila_param[3] is its recursion cutoff). */
int lila_dgeqrf_w03_levelx(int *lila_param, int n_lvl, int i_lvl, int *nb_lvl, int m, int n, int i,
                           int mt, double *A, int lda, double *T, int ldt, double *Q, int ldq,
                           double *work, int lwork);
int lila_dgeqrf_w03_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                              double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int lila_dgeqr2_w03(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                    int ldt, double *Q, int ldq, double *work, int lwork);
int lila_dgeqrf_w03_mt(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                       int ldt, double *Q, int ldq, double *work, int lwork);
int lila_dgeqrf_w03_mt_hh(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                          double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int lila_dgeqrf_w03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                         double *Q, int ldq, double *work, int lwork);
int lila_dgeqrf_w03_mt_hr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                          double *Q, int ldq, double *work, int lwork);
int lila_dgeqr2_w03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                      double *Q, int ldq, double *work, int lwork);
int lila_dgeqr2_w03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                      double *Q, int ldq, double *work, int lwork);
int lila_dgeqr2_w03_hr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                       double *Q, int ldq, double *work, int lwork);

/* lila-c/lila.h:58-72 -- workspace queries (doubles), the reference's
 * closed forms; lwork carries the running maximum as in the reference */
int lila_query_dgeqrf_w03_levelx(int *lila_param, int n_lvl, int i_lvl, int *nb_lvl, int m, int n,
                                 int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                                 int ldq, double *work, int lwork);
int lila_query_dgeqrf_w03_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                    double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int lila_query_dgeqrf_ker_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                    double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int lila_query_dgeqrf_w03_mt(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                             double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int lila_query_dgeqrf_w03_mt_hh(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                                double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int lila_query_dormqrbz_w03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *Q,
                            int ldq, double *T, int ldt, double *work, int lwork);
int lila_query_dgeqrf_w03_mt_hr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                                double *Q, int ldq, double *work, int lwork);
int lila_query_dgeqrf_w03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                               double *Q, int ldq, double *work, int lwork);
int lila_query_dgeqr2_w03_hr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                             double *Q, int ldq, double *work, int lwork);
int lila_query_dgeqr2_w03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                            double *Q, int ldq, double *work, int lwork);
int lila_query_dgeqr2_w03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                            double *Q, int ldq, double *work, int lwork);
int lila_query_ormhr2_w03_hr(int m, int n, int i, int j, int l, int mt, double *A, int lda, double *T,
                             int ldt, double *Q, int ldq, double *work, int lwork, double *S);
int lila_query_dorgh2_3(int m, int n, double *A, int lda, double *T, int ldt, double *Q, int ldq,
                        double *work, int lwork, double *S);
int lila_query_dormqrf_w03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T,
                           int ldt, double *work, int lwork);
int lila_query_dgeqrf_LAPACK_appendcols(int m, int k, int n, int mt, double *A, int lda, double *Q,
                                        int ldq, double *T, int ldt, double *work, int lwork);

/* lila-c/lila.h:19, :21, :22 -- Q^T applied with LAPACK dormqr semantics:
 * w00/z00 rebuild tau = 2/(1 + ||v||^2) from the stored vectors (C = A's
 * columns j.. or B); z02 applies one block with the full-layout T at T(i,i).
 * Note z00's order (A, B, T) and z02's (A, T, B), as in the reference. */
int lila_dormqrf_w00(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T,
                     int ldt, double *work, int lwork);
int lila_dormqrf_z00(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *B,
                     int ldb, double *T, int ldt, double *work, int lwork);
int lila_dormqrf_z02(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T,
                     int ldt, double *B, int ldb, double *work, int lwork);

/* lila-c/lila.h:74-81 -- V, R and T only (no Q); recursion leaf when n < nx */
int lila_dgeqrf_v03_levelx(int *lila_param, int n_lvl, int i_lvl, int *nb_lvl, int m, int n, int i,
                           int mt, double *A, int lda, double *T, int ldt, double *work, int lwork);
int lila_dgeqrf_v03_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                              double *T, int ldt, double *work, int lwork);
int lila_dgeqrf_v03_mt(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                       int ldt, double *work, int lwork);
int lila_dgeqr2_v03(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                    int ldt, double *work, int lwork);
int lila_dgeqrf_v03_mt_hh(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                          double *T, int ldt, double *work, int lwork);
int lila_dgeqrf_v03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                         double *work, int lwork);
int lila_dgeqr2_v03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                      double *work, int lwork);
int lila_dgeqr2_v03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                      double *work, int lwork);

/* lila-c/lila.h:83-90 -- explicit Q from the stored (V, T) of a v03 pass */
int lila_dgeqrf_q03_levelx(int *lila_param, int n_lvl, int i_lvl, int *nb_lvl, int m, int n, int i,
                           int mt, double *A, int lda, double *T, int ldt, double *Q, int ldq,
                           double *work, int lwork);
int lila_dgeqrf_q03_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                              double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int lila_dgeqrf_q03_mt_hh(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                          double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int lila_dgeqrf_q03_mt(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                       int ldt, double *Q, int ldq, double *work, int lwork);
int lila_dgeqrf_q03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                         double *Q, int ldq, double *work, int lwork);
int lila_dgeqr2_q03(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                    int ldt, double *Q, int ldq, double *work, int lwork);
int lila_dgeqr2_q03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                      double *Q, int ldq, double *work, int lwork);
int lila_dgeqr2_q03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                      double *Q, int ldq, double *work, int lwork);

/* lila-c/lila.h:95 -- recursive QR on LAPACK kernels; T holds tau (T + i is
 * the tau of the columns from i) */
int lila_dgeqrf_ker_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                              double *T, int ldt, double *Q, int ldq, double *work, int lwork);

/* lila-c/lila.h:97-99 -- the recursion with lila_param[4] choosing VRTQ (0),
 * VRT (1) or Q from (V, T) (2) */
int lila_dgeqrf_qr2_recursive2(int *lila_param, int m, int n, int i, int mt, int nb1, int nb2,
                               double *A, int lda, double *T, int ldt, double *Q, int ldq,
                               double *work, int lwork);
int lila_dgeqrf_qr2_recursive(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                              double *T, int ldt, double *Q, int ldq, double *work, int lwork);
int lila_dgeqrf_qr2(int *lila_param, int m, int n, int i, int mt, double *A, int lda, double *T,
                    int ldt, double *Q, int ldq, double *work, int lwork);

/* lila-c/lila.h:101-102 -- QR of k factored columns extended by n new ones
 * (Block-Arnoldi); the LAPACK form keeps tau (k + n entries) in T */
int lila_dgeqrf_w03_appendcols(int *lila_param, int m, int k, int n, int mt, double *A, int lda,
                               double *Q, int ldq, double *T, int ldt, double *work, int lwork);
int lila_dgeqrf_LAPACK_appendcols(int m, int k, int n, int mt, double *A, int lda, double *Q,
                                  int ldq, double *T, int ldt, double *work, int lwork);

/* lila-c/lila.h:92-93 */
int lila_dT2tau_w03(int m, int n, int i, int mt, double *T, int ldt, double *tau);
int lila_dV2tau_w03(int m, int n, int i, int mt, double *A, int lda, double *tau);

/* ---- the recursive QR library of qr3/src (its header qr3/src/qr2.h) ---- */

/* qr3/src/qr2.h:20 (qr3/src/qr2_dgeqr3_R.c): recursive QR with R apart --
 * V below A's diagonal, T (n x n upper), R (n x n upper).  A's diagonal and
 * upper part are not an output (the reference leaves there each entry as it
 * stood before its column's reflector; here they keep their input values);
 * with R == A the result is dgeqr3's. */
int qr2_dgeqr3_R(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr);
/* qr3/src/qr2.h:30 (qr3/src/qr2_dlarft3.c): T from the reflectors below A's
 * diagonal and tau */
int qr2_dlarft3(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
/* qr3/src/qr2.h:25 (qr3/src/qr2_dorgqr3.c): explicit m x n Q from the
 * reflectors below Q's diagonal and T, in place (T's diagonal equals tau) */
int qr2_dorgqr3(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau);
/* qr3/src/qr2.h:27 (qr3/src/qr2_dorgqr3_VT2Q.c): the same Q */
int qr2_dorgqr3_VT2Q(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau);

/* UT forms: T replaced by S = T^-1 (strict upper part of V^T V, diagonal
 * 1/tau).  qr3/src/qr2.h:21 (qr3/src/qr2_dgeqr3_R_UT.c): V and R as
 * qr2_dgeqr3_R, S in T */
int qr2_dgeqr3_R_UT(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr);
/* ISW forms: T (or S) only on the diagonal blocks of the halving
 * recursion's right spine (sizes n/2, (n - n/2)/2, ..., 1); the blocks
 * between them are not written.  qr3/src/qr2.h:22-23
 * (qr2_dgeqr3_R_ISW.c, qr2_dgeqr3_R_UT_ISW.c) */
int qr2_dgeqr3_R_ISW(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr);
int qr2_dgeqr3_R_UT_ISW(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr);
/* qr3/src/qr2.h:26 (qr2_dorgqr3_UT.c) and :28 (qr2_dorgqr3_VS2Q.c): Q from
 * the reflectors below Q's diagonal and S */
int qr2_dorgqr3_UT(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau);
int qr2_dorgqr3_VS2Q(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau);
/* qr3/src/qr2.h:31-35 (qr2_dlarft3_UT.c, _ISW.c, _UT_ISW.c, _ISW_V2T.c,
 * _ISW_V2T_UT.c): S, or T / S on the spine blocks, from V and tau */
int qr2_dlarft3_UT(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int qr2_dlarft3_ISW(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int qr2_dlarft3_UT_ISW(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int qr2_dlarft3_ISW_V2T(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int qr2_dlarft3_ISW_V2T_UT(int m, int n, double *A, int lda, double *T, int ldt, double *tau);

/* aux routines of qr3/src (qr3/src/qr2.h:37-47) */
/* :37 (qr2_aux_dV2N.c): upper(T) <- U U^T, U unit upper with T's strict
 * upper part (V1^T V1 when that part holds V1^T) */
int qr2_aux_dV2N(int n, double *T, int ldt);
/* :38 (qr2_aux_dN2T.c): T from tau and the strict upper part of V^T V held
 * in T (the dlarft recurrence) */
int qr2_aux_dN2T(int n, double *tau, double *T, int ldt);
/* :39-40 (qr2_aux_dVS2Q.c, qr2_aux_dVT2Q.c): Q (m x n) holding V below its
 * diagonal and S (or T) on and above becomes [I;0] - V T V1^T */
int qr2_aux_dVS2Q(int m, int n, double *Q, int ldq);
int qr2_aux_dVT2Q(int m, int n, double *Q, int ldq);
/* :41 (qr2_aux_dmLUinA.c): A <- -L U, L unit lower and U upper held in A */
int qr2_aux_dmLUinA(int n, double *A, int lda);
/* :42 (qr2_aux_dV2tau.c): tau_k = 2 / (1 + ||A(k+1:m, k)||^2); tau_{n-1} = 0
 * when m == n */
int qr2_aux_dV2tau(int m, int n, double *A, int lda, double *tau);
/* :43 (qr2_aux_dULTinU.c): upper(U) <- U L^T, L unit lower (may share U's
 * storage) */
int qr2_aux_dULTinU(int n, double *L, int ldl, double *U, int ldu);
/* :44 (qr2_aux_dApUBTinA.c): A (m x n) += U B^T, U m x m upper, B n x m */
int qr2_aux_dApUBTinA(int m, int n, double *A, int lda, double *U, int ldu, double *B, int ldb);
/* :45 (qr2_aux_dUinvLTinU.c): upper(U) <- U^-1 L^T, L unit lower */
int qr2_aux_dUinvLTinU(int n, double *L, int ldl, double *U, int ldu);
/* :47 (qr2_aux_dlarf_wrapper.c, LAPACK dlarf): C <- (I - tau v v^T) C for
 * side 'L' (v of length m), C (I - tau v v^T) for 'R' (length n); v with
 * stride incv >= 1 (-1 otherwise); work is not used */
int qr2_aux_dlarf_wrapper(char side, int m, int n, double *V, int incv, double tau, double *C,
                          int ldc, double *work);

/* ---- operation counts of the recursive QR library (qr3/flops/flops.h) -----
 * Host arithmetic (lila_amd/csrc/flops.c), the reference's formulas: a
 * closed form per routine and, for the recursive ones, the `_check` sum over
 * the recursion's steps.  qr3/flops/flops.h:7-10 */
long int flops_gemm(int m, int n, int k);
long int flops_trmm(char S, int m, int n);
long int flops_syrk(int n, int k);
long int flops_larft(int m, int k);
/* qr3/flops/flops.h:13-25 */
long int flops_mLUinA(int n);
long int flops_V2N(int n);
long int flops_ApUBTinA(int m, int n);
long int flops_ULTinU(int n);
long int flops_N2T(int n);
long int flops_VT2Q(int m, int n);
long int flops_mLUinA_check(int n);
long int flops_V2N_check(int n);
long int flops_ApUBTinA_check(int m, int n);
long int flops_ULTinU_check(int n);
long int flops_N2T_check(int n);
long int flops_VT2Q_check(int m, int n);
/* qr3/flops/flops.h:30-41 */
long int flops_lapack_larf(int m, int n);
long int flops_lapack_larfg(int m);
long int flops_lapack_geqr2(int m, int n);
long int flops_lapack_org2r_check(int m, int n, int k);
long int flops_lapack_org2r(int m, int n, int k);
/* qr3/flops/flops.h:62-63, 87-91 */
long int flops_geqr3(int m, int n);
long int flops_geqr3_check(int m, int n);
long int flops_qr2_dorgqr_check(int m, int n);
long int flops_qr2_dorgqr(int m, int n);
long int flops_dlarft3_check(int m, int n);
long int flops_dlarft3(int m, int n);

/* ---- extensions (no counterpart in the reference header) ------------- */

/* Status of the last call (0 ok, >0 numerical info, <0 error), after
 * waiting for the library stream; clears it. */
int lila_get_info(void);
/* Wait for the library stream; returns the status like lila_get_info. */
int lila_sync(void);
/* Library stream (a hipStream_t); NULL selects the default stream.  The
 * switch is ordered: work enqueued afterwards waits for the library's work
 * already enqueued on the previous stream. */
void lila_set_stream(void *hip_stream);
void *lila_get_stream(void);

/* Least squares through dgeqr3, Q^T B from (V,T) and the R solve, the
 * sequence of least-squares/main_ls_qr3.c:94-119.  A (m x n) is overwritten
 * by V and R, T (n x n) by the T factor, B (m x k) by Q^T B with X = its
 * first n rows. */
int lila_dgels_qr3(int m, int n, int k, double *A, int lda, double *T, int ldt, double *B, int ldb);

/* Batched least squares over nprob independent problems of uneven sizes:
 * problem p has A_p (m[p] x n[p], lda[p]) and B_p (m[p] x nrhs[p], ldb[p]).
 * On exit, as LAPACK dgels('N') leaves them: A_p holds R on and above the
 * diagonal and the Householder vectors (dlarfg convention, unit diagonal
 * implied) below it; rows [0, n) of B_p hold the solution X, rows [n, m)
 * the rest of Q^T B (their column norms are the residual norms).  The
 * per-problem sequence is that of least-squares/main_ls_qr3.c:94-119
 * (factor, apply Q^T, back-substitute R); the batch runs as one persistent
 * launch with a work queue ordered by flop count.
 * Limits: m >= n >= 1, n <= 128, 0 <= nrhs <= 16; returns -1 (with a
 * message naming the problem) otherwise.  All-device pointers run
 * asynchronously on the library stream; all-host pointers are staged. */
int lila_dgels_batched(int nprob, const int *m, const int *n, const int *nrhs, double *const *A,
                       const int *lda, double *const *B, const int *ldb);
/* The same as a reusable plan over device pointers (the problem table is
 * uploaded once): run(use_graph = 1) replays a captured hipGraph of the
 * launch, run(0) launches it directly; flops() is the plan's flop count
 * (2mn^2 - n^3/3 + 4(m-n)nk + 4n^2 k summed over the problems, the count
 * least-squares/main_ls_qr3.c:125-129 reports). */
typedef struct ld_batch_plan lila_batch_plan;
lila_batch_plan *lila_dgels_batched_plan(int nprob, const int *m, const int *n, const int *nrhs,
                                         double *const *A, const int *lda, double *const *B,
                                         const int *ldb);
int lila_dgels_batched_run(lila_batch_plan *plan, int use_graph);
/* No counterpart in the reference: lila_dgels_batched for callers that
 * cannot allocate (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The problem table, queue order and
 * queue counter live in a static device array of the library
 * (16384 problems at most, ~0.7 MB per GPU, held until the process ends)
 * instead of memory allocated per call; A[p] and B[p] must be device
 * pointers.  Enqueued on the library stream and returns (0; -1 for more
 * than 16384 problems, sizes out of range or host pointers, with nothing
 * launched).  Calls share the array of their GPU: two calls that overlap
 * on one GPU (different streams, or threads) race. */
int lila_dgels_batched_noalloc(int nprob, const int *m, const int *n, const int *nrhs,
                               double *const *A, const int *lda, double *const *B, const int *ldb);
/* Caps the workgroups of the persistent launch for plans made from now on
 * (0: one per CU at the kernel's occupancy, the default); returns the
 * previous cap.  A workgroup takes problems from the queue until it is
 * empty, so any cap >= 1 gives the same results. */
int lila_set_batched_grid(int wgs);
double lila_dgels_batched_flops(const lila_batch_plan *plan);
void lila_dgels_batched_destroy(lila_batch_plan *plan);

/* C = alpha op(A) op(B) + beta C on FP64 MFMA, dgemm argument order
 * (transa/transb 'N' or 'T'); the GEMM every BLAS-3 step above runs on. */
int lila_dgemm(char transa, char transb, int m, int n, int k, double alpha, const double *A,
               int lda, const double *B, int ldb, double beta, double *C, int ldc);
/* No counterpart in the reference: the same product emulated on int8 MFMA
 * by the Chinese-remainder scheme (Ozaki scheme II): rows of op(A) and
 * columns of B scaled to kb-bit integers (8 <= kb <= 53), their residues
 * modulo 14-16 coprime moduli <= 256 multiplied exactly in int32, the
 * integer product recovered by the CRT in FP64 and scaled back.  With
 * kb = 53 the error is that of an FP64 product norm-wise (relative 2^-53 of
 * each row's and column's largest entry).  transb must be 'N', k <= 65536;
 * finite entries only.  Returns -1 otherwise. */
int lila_dgemm_emu(char transa, char transb, int m, int n, int k, double alpha, const double *A,
                   int lda, const double *B, int ldb, double beta, double *C, int ldc, int kb);

/* Autotuner (no counterpart in the reference, whose mt, nx and block sizes
 * are set by hand on its test drivers' command lines): times
 * lila_dgeqrf_w03_recursive (VRTQ, FP64, panel dgeqr3, leaf hh) on an m x n
 * Uniform(-1,1) device matrix and, one setting at a time from the current
 * ones (best of `reps` runs each), picks the T block height mt from
 * {256, 512, 1024} and the recursion cutoff nx from {32, 64, ..., 2048} (both
 * in/out: read as the starting point, written with the choice, which the
 * caller then passes as mt and lila_param[3]), then the GEMM tile shapes,
 * split-K occupancy target, small-output tile thresholds, panel slab
 * height, panel width nb (dgeqr3's recursion cutoff, from {32, 16, 24}),
 * panel algorithm and where Q is built (lila_set_q_after), which the library
 * keeps for the rest of the process.  `report`
 * (may be NULL) receives the trial times and the choice.  Allocates 3 m x n
 * matrices and a 1024 x n T on the device for the duration.  Returns 0, or
 * -1 without a device, on bad sizes or when the allocation fails. */
int lila_autotune(int m, int n, int reps, int *mt, int *nx, char *report, int len);
/* Selects the form lila_autotune times: 0 the plain FP64 path (default), 1
 * the mixed-precision form (lila_param[0] = 1); returns the previous one. */
int lila_autotune_form(int form);

/* Panel algorithm of the drivers (returns the previous one): 1 (default)
 * Cholesky-QR2 + Householder reconstruction in one fused launch, with the
 * column-by-column kernel as in-kernel fallback; 0 always column-by-column
 * (dgeqr3's arithmetic).  Both give dgeqr3's V, R, T and Q to rounding. */
int lila_set_panel_mode(int mode);
/* Panel width nb (8..32 or 64, default 64): dgeqr3's recursion hands widths
 * up to nb to one panel launch; with 64, widths 33..64 go to the 64-column
 * Cholesky-QR2 panel where its LDS slabs hold the panel (up to ~20700 rows;
 * taller ones recurse down to 32-wide panels).  Returns the previous width,
 * -1 out of range. */
int lila_set_panel_width(int nb);
/* VRTQ recursion (lila_dgeqrf_w03_recursive with hh or LAPACK leaves): 0
 * (default) builds Q at every node as the reference's call tree does, 1
 * after the VRT recursion by whole T blocks from the last to the first (the
 * same Q to rounding).  Returns the previous setting. */
int lila_set_q_after(int on);
/* dgeqr3 nodes whose halves are at most 64 columns wide: 1 (default) the
 * update of the right half and the T join each in one persistent launch
 * (partial products per workgroup, a fixed-order sum, the 64 x 64 work); 0
 * as GEMM launches.  The same V, R, T to rounding.  Returns the previous
 * setting.  (An extension: the reference has no counterpart.) */
int lila_set_node_fused(int on);
/* Backend of the large plain products (no structure, op(A) = A, op(B) = B:
 * the general interiors of the trailing updates C -= V W and of lila_dgemm):
 * 0 (default) the library's own direct-to-LDS FP64 MFMA kernel, 1 rocBLAS
 * dgemm on the library stream.  Structured, transposed and small
 * products always run on the library's kernels.  Returns the previous
 * setting, -1 out of range.  (An extension: the reference has no
 * counterpart.) */
int lila_set_gemm_backend(int b);

/* Cholesky-QR of the hr paths (lila_dgeqrf_w03_mt_hr, lila_dgeqr2_w03_hr
 * and the hr leaves; returns the previous setting, -1 outside 0..3): 0
 * (default) one pass with an FP64 Gram, as the reference; 1, 2, 3 two passes
 * (Cholesky-QR2) whose first Gram is FP64 MFMA, FP32 MFMA (operands rounded
 * to FP32) or 3-term split-BF16 MFMA with FP32 accumulation (n <= 256), the
 * second FP64.  A Cholesky breakdown (the first Gram not positive definite
 * at its precision) returns the column as a positive info. */
int lila_set_gram_precision(int prec);

/* G (n x n; only the upper triangle is written) = A^T A of the m x n A in
 * the arithmetic of prec: 0 or 1 FP64 MFMA, 2 FP32 MFMA, 3 split BF16
 * (n <= 256).  The Gram of the Cholesky-QR paths, exposed for measurement. */
int lila_dgram(int prec, int m, int n, double *A, int lda, double *G, int ldg);

/* Streams of lila_dgeqrf_w03_recursive (returns the previous count): 1
 * (default) keeps every step on the library stream; 2 builds Q (the
 * dormqrbz steps) on a second, internal stream, overlapped with the
 * factorization on the library stream and joined back before the call's
 * work on the library stream ends (about 2% faster at m = n = 16384; the
 * overlapped launches then share the GPU, so per-launch timings stretch).
 * The environment variable LILA_LANES=2 sets 2 at initialisation.  Results
 * agree to rounding either way. */
int lila_set_lanes(int n);

/* Arithmetic of the library's GEMMs (returns the previous setting): 0
 * (default) FP64 MFMA; 1 FP32 MFMA with FP32 accumulation, operands rounded
 * to FP32 (the low-precision updates of the mixed-precision QR, selected by
 * lila_param[0] == 1 in lila_dgeqrf_w03_recursive; exposed for testing). */
int lila_set_gemm_lowp(int on);

/* Launch profiling: when on, every launch of a hot kernel is bracketed by
 * hipEvents on the library stream; lila_prof_read writes one line per
 * kernel class, "name|launches|algorithmic_flops|algorithmic_bytes|ms". */
int lila_prof_enable(int on);
int lila_prof_read(char *buf, int buflen);
/* Record only launches of this kernel class (its exact name as printed by
 * lila_prof_read); NULL or "" records every class.  Each recorded launch
 * adds two event records to the stream, so timing runs filter to the class
 * they report. */
int lila_prof_only(const char *kernel_class);

/* Device generators of the test inputs (splitmix64 stream; Box-Muller). */
int lila_fill_uniform(unsigned long long seed, int m, int n, double *A, int lda, double lo,
                      double hi);
int lila_fill_normal(unsigned long long seed, int m, int n, double *A, int lda);

#ifdef __cplusplus
}
#endif
#endif
