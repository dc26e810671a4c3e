/* lila_drivers.h -- device-pointer drivers (drivers.c) used by api.c */
#ifndef LILA_DRIVERS_H
#define LILA_DRIVERS_H

int dv_dormqrf_w03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T, int ldt);
int dv_dormqrf_z03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *T, int ldt,
                   double *B, int ldb);
int dv_dormqrbz_w03(int m, int n, int k, int i, int j, int mt, double *A, int lda, double *Q, int ldq,
                    double *T, int ldt);
int dv_dlarft_connect_w03(int m, int n, int i, int j, int mt, double *A, int lda, double *T, int ldt);
int dv_dgeqr3(int m, int n, double *A, int lda, double *T, int ldt);
int dv_dgeqr2_w03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                    int ldq);
int dv_dgeqr2_w03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                    int ldq);
int dv_dgeqr2_w03_hr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                     int ldq);
int dv_dgeqr2_w03(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                  double *Q, int ldq);
int dv_dorgh2_3(int m, int n, double *A, int lda, double *T, int ldt, double *Q, int ldq, double *S);
int dv_dorgh2(int m, int n, int i, int j, int mt, double *A, int lda, double *T, int ldt, double *Q,
              int ldq, double *S);
int dv_ormhr2_w03_hr(int m, int n, int i, int j, int l, int mt, double *A, int lda, double *T, int ldt,
                     double *Q, int ldq, double *S);
int dv_dorghr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q, int ldq,
              double *Sd);
int dv_dgeqrf_w03_mt_hr(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                        int ldq);
int dv_dlarft_w03(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, const double *tau);
int dv_dgeqrf_w03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                       int ldq);
int dv_dgeqrf_w03_mt_hh(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                        int ldt, double *Q, int ldq);
int dv_dgeqrf_w03_mt(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                     double *Q, int ldq);
int dv_dgeqrf_w03_recursive(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                            int ldt, double *Q, int ldq);
int dv_dgeqrf_w03_levelx(const int *p, int n_lvl, int i_lvl, const int *nb_lvl, int m, int n, int i,
                         int mt, double *A, int lda, double *T, int ldt, double *Q, int ldq);
int dv_dgetrf_b(int m, int n, int i, int mt, double *A, int lda, double *Q, int ldq, double *U);
int dv_ls_qr3(int m, int n, int k, double *A, int lda, double *T, int ldt, double *B, int ldb);

int dv_dgeqr2_v03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt);
int dv_dgeqr2_v03(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt);
int dv_dgeqrf_v03_mt_hh(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                        int ldt);
int dv_dgeqrf_v03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt);
int dv_dgeqrf_v03_mt(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                     int ldt);
int dv_dgeqrf_v03_recursive(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                            int ldt);
int dv_dgeqrf_v03_levelx(const int *p, int n_lvl, int i_lvl, const int *nb_lvl, int m, int n, int i,
                         int mt, double *A, int lda, double *T, int ldt);
int dv_dgeqr2_q03_3(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                    int ldq);
int dv_dgeqr2_q03_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                    int ldq);
int dv_dgeqr2_q03(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                  double *Q, int ldq);
int dv_dgeqrf_q03_mt_hh(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                        int ldt, double *Q, int ldq);
int dv_dgeqrf_q03_mt_l(int m, int n, int i, int mt, double *A, int lda, double *T, int ldt, double *Q,
                       int ldq);
int dv_dgeqrf_q03_mt(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                     double *Q, int ldq);
int dv_dgeqrf_q03_recursive(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                            int ldt, double *Q, int ldq);
int dv_dgeqrf_q03_levelx(const int *p, int n_lvl, int i_lvl, const int *nb_lvl, int m, int n, int i,
                         int mt, double *A, int lda, double *T, int ldt, double *Q, int ldq);
int dv_dgeqrf_w03_appendcols(const int *p, int m, int k, int n, int mt, double *A, int lda, double *Q,
                             int ldq, double *T, int ldt);
int dv_dgeqrf_LAPACK_appendcols(int m, int k, int n, double *A, int lda, double *Q, int ldq,
                                double *tau);
int dv_dgeqrf_ker_recursive(const int *p, int m, int n, int i, double *A, int lda, double *T, double *Q,
                            int ldq);
int dv_dormqrf_tau_from_v(int m, int n, int k, int i, const double *A, int lda, double *C, int ldc);
int dv_dormqrf_z02(int m, int n, int k, int i, int j, double *A, int lda, double *T, int ldt, double *B,
                   int ldb);
int dv_dgeqrf_qr2(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                  double *Q, int ldq);
int dv_dgeqrf_qr2_recursive(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T,
                            int ldt, double *Q, int ldq);
int dv_dgeqrf_qr2_recursive2(const int *p, int m, int n, int i, int mt, int nb1, int nb2, double *A,
                             int lda, double *T, int ldt, double *Q, int ldq);

int dv_dgeqrf_mixed(const int *p, int m, int n, int i, int mt, double *A, int lda, double *T, int ldt,
                    double *Q, int ldq);

int dq_dormqrf_w03(int m, int n, int k, int i, int mt);
int dq_dormqrbz_w03(int m, int n, int k, int i, int mt);
int dq_dgeqrf_w03_mt(const int *p, int m, int n, int i, int mt, int lwork);
int dq_dgeqrf_w03_recursive(const int *p, int m, int n, int i, int mt, int lwork);
int dq_dgeqrf_w03_levelx(const int *p, int n_lvl, int i_lvl, const int *nb_lvl, int m, int n, int i,
                         int mt);

int dv_qr2_dgeqr3_R(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr);
int dv_qr2_dlarft3(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int dv_qr2_dorgqr3(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau);
int dv_gram(int prec, int ml, int n, const double *X, int ldx, double *G, int ldg);
int dv_qr2_dgeqr3_R_UT(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr);
int dv_qr2_dgeqr3_R_isw(int ut, int m, int n, double *A, int lda, double *T, int ldt, double *R,
                        int ldr);
int dv_qr2_dlarft3_UT(int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int dv_qr2_dlarft3_isw(int ut, int m, int n, double *A, int lda, double *T, int ldt, double *tau);
int dv_qr2_dorgqr3_UT(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau);
int dv_qr2_aux_dVT2Q(int m, int n, double *Q, int ldq);
int dv_qr2_aux_dVS2Q(int m, int n, double *Q, int ldq);
int dv_qr2_aux_dmLUinA(int n, double *A, int lda);
int dv_qr2_aux_dULTinU(int n, double *L, int ldl, double *U, int ldu);
int dv_qr2_aux_dUinvLTinU(int n, double *L, int ldl, double *U, int ldu);
int dv_qr2_aux_dApUBTinA(int m, int n, double *A, int lda, double *U, int ldu, double *B, int ldb);
int dv_qr2_aux_dV2N(int n, double *T, int ldt);
int dv_qr2_aux_dN2T(int n, double *tau, double *T, int ldt);
int dv_qr2_aux_dV2tau(int m, int n, double *A, int lda, double *tau);
int dv_qr2_aux_dlarf(char side, int m, int n, const double *V, int incv, double tau, double *C,
                     int ldc);

#endif
