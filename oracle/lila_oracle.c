/*
 * lila_oracle.c -- CPU oracle for the LILA Householder-QR hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * lila_oracle.h for the contract.  All matrices are column-major FP64.
 *
 * Reference citations are relative to the reference tree root.  The w03
 * routines follow lila-c/ (the canonical snapshot); the reconstruction
 * routines follow lila-c/ and storage/c2/ as cited per function.
 */
#include "lila_oracle.h"

#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define IDX(M, r, c, ld) ((M)[(size_t)(r) + (size_t)(c) * (size_t)(ld)])

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "lila_oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}

/* ======================================================================== */
/* Built-in BLAS subset (naive loops)                                        */
/* ======================================================================== */

static void b_dgemm(char ta, char tb, int m, int n, int k, double alpha,
                    const double *A, int lda, const double *B, int ldb,
                    double beta, double *C, int ldc) {
    int nta = (ta == 'N' || ta == 'n');
    int ntb = (tb == 'N' || tb == 'n');
    if (m <= 0 || n <= 0) return;
    for (int j = 0; j < n; j++) {
        double *c = &IDX(C, 0, j, ldc);
        if (beta == 0.0) {
            for (int i = 0; i < m; i++) c[i] = 0.0;
        } else if (beta != 1.0) {
            for (int i = 0; i < m; i++) c[i] *= beta;
        }
        if (alpha == 0.0 || k <= 0) continue;
        if (nta) {
            for (int p = 0; p < k; p++) {
                double b = ntb ? IDX(B, p, j, ldb) : IDX(B, j, p, ldb);
                double ab = alpha * b;
                if (ab == 0.0) continue;
                const double *a = &IDX(A, 0, p, lda);
                for (int i = 0; i < m; i++) c[i] += ab * a[i];
            }
        } else {
            for (int i = 0; i < m; i++) {
                const double *a = &IDX(A, 0, i, lda);
                double s = 0.0;
                if (ntb) {
                    const double *b = &IDX(B, 0, j, ldb);
                    for (int p = 0; p < k; p++) s += a[p] * b[p];
                } else {
                    for (int p = 0; p < k; p++) s += a[p] * IDX(B, j, p, ldb);
                }
                c[i] += alpha * s;
            }
        }
    }
}

/* Dense copy of op(A) for a triangular A (order nn), zeros outside the
 * triangle and ones on the diagonal when diag == 'U'. */
static double *tri_dense(char uplo, char trans, char diag, int nn,
                         const double *A, int lda) {
    double *D = (double *)xmalloc((size_t)nn * nn * sizeof(double));
    int up = (uplo == 'U' || uplo == 'u');
    int tr = (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c');
    int unit = (diag == 'U' || diag == 'u');
    for (int c = 0; c < nn; c++)
        for (int r = 0; r < nn; r++) {
            int inside = up ? (r <= c) : (r >= c);
            double v = inside ? IDX(A, r, c, lda) : 0.0;
            if (r == c && unit) v = 1.0;
            if (tr) IDX(D, c, r, nn) = v; else IDX(D, r, c, nn) = v;
        }
    return D;
}

static void b_dtrmm(char side, char uplo, char trans, char diag, int m, int n,
                    double alpha, const double *A, int lda, double *B, int ldb) {
    if (m <= 0 || n <= 0) return;
    int left = (side == 'L' || side == 'l');
    int nn = left ? m : n;
    double *D = tri_dense(uplo, trans, diag, nn, A, lda);
    double *X = (double *)xmalloc((size_t)m * n * sizeof(double));
    if (left) b_dgemm('N', 'N', m, n, m, alpha, D, nn, B, ldb, 0.0, X, m);
    else      b_dgemm('N', 'N', m, n, n, alpha, B, ldb, D, nn, 0.0, X, m);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) IDX(B, i, j, ldb) = IDX(X, i, j, m);
    free(X);
    free(D);
}

static void b_dtrsm(char side, char uplo, char trans, char diag, int m, int n,
                    double alpha, const double *A, int lda, double *B, int ldb) {
    if (m <= 0 || n <= 0) return;
    int left = (side == 'L' || side == 'l');
    int nn = left ? m : n;
    double *D = tri_dense(uplo, trans, diag, nn, A, lda);
    /* op(A) is lower when exactly one of (uplo == L, trans == T) holds */
    int up = (uplo == 'U' || uplo == 'u');
    int tr = (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c');
    int lower = (up == tr);
    if (left) {
        /* solve D X = alpha B, column by column */
        for (int j = 0; j < n; j++) {
            double *b = &IDX(B, 0, j, ldb);
            for (int i = 0; i < m; i++) b[i] *= alpha;
            if (lower) {
                for (int i = 0; i < m; i++) {
                    double s = b[i];
                    for (int p = 0; p < i; p++) s -= IDX(D, i, p, nn) * b[p];
                    b[i] = s / IDX(D, i, i, nn);
                }
            } else {
                for (int i = m - 1; i >= 0; i--) {
                    double s = b[i];
                    for (int p = i + 1; p < m; p++) s -= IDX(D, i, p, nn) * b[p];
                    b[i] = s / IDX(D, i, i, nn);
                }
            }
        }
    } else {
        /* solve X D = alpha B, row by row: column c of X depends on the
         * columns p with D(p,c) != 0, p != c */
        for (int r = 0; r < m; r++) {
            for (int c = 0; c < n; c++) IDX(B, r, c, ldb) *= alpha;
            if (lower) {
                /* X(r,c) D(c,c) = B(r,c) - sum_{p>c} X(r,p) D(p,c) */
                for (int c = n - 1; c >= 0; c--) {
                    double s = IDX(B, r, c, ldb);
                    for (int p = c + 1; p < n; p++) s -= IDX(B, r, p, ldb) * IDX(D, p, c, nn);
                    IDX(B, r, c, ldb) = s / IDX(D, c, c, nn);
                }
            } else {
                for (int c = 0; c < n; c++) {
                    double s = IDX(B, r, c, ldb);
                    for (int p = 0; p < c; p++) s -= IDX(B, r, p, ldb) * IDX(D, p, c, nn);
                    IDX(B, r, c, ldb) = s / IDX(D, c, c, nn);
                }
            }
        }
    }
    free(D);
}

static void b_dsyrk(char uplo, char trans, int n, int k, double alpha,
                    const double *A, int lda, double beta, double *C, int ldc) {
    int up = (uplo == 'U' || uplo == 'u');
    int tr = (trans == 'T' || trans == 't');
    for (int j = 0; j < n; j++) {
        int r0 = up ? 0 : j, r1 = up ? j : n - 1;
        for (int i = r0; i <= r1; i++) {
            double s = 0.0;
            if (tr) {
                for (int p = 0; p < k; p++) s += IDX(A, p, i, lda) * IDX(A, p, j, lda);
            } else {
                for (int p = 0; p < k; p++) s += IDX(A, i, p, lda) * IDX(A, j, p, lda);
            }
            double c = (beta == 0.0) ? 0.0 : beta * IDX(C, i, j, ldc);
            IDX(C, i, j, ldc) = c + alpha * s;
        }
    }
}

static void b_dger(int m, int n, double alpha, const double *x, int incx,
                   const double *y, int incy, double *A, int lda) {
    for (int j = 0; j < n; j++) {
        double t = alpha * y[(size_t)j * incy];
        if (t == 0.0) continue;
        for (int i = 0; i < m; i++) IDX(A, i, j, lda) += x[(size_t)i * incx] * t;
    }
}

/* ---- optional host BLAS (Fortran interface, hidden string lengths) ----- */
typedef void (*f_gemm)(const char *, const char *, const int *, const int *,
                       const int *, const double *, const double *, const int *,
                       const double *, const int *, const double *, double *,
                       const int *, size_t, size_t);
typedef void (*f_trxm)(const char *, const char *, const char *, const char *,
                       const int *, const int *, const double *, const double *,
                       const int *, double *, const int *, size_t, size_t,
                       size_t, size_t);
typedef void (*f_syrk)(const char *, const char *, const int *, const int *,
                       const double *, const double *, const int *,
                       const double *, double *, const int *, size_t, size_t);
typedef void (*f_ger)(const int *, const int *, const double *, const double *,
                      const int *, const double *, const int *, double *,
                      const int *);

static struct {
    void *lib;
    f_gemm gemm;
    f_trxm trmm, trsm;
    f_syrk syrk;
    f_ger ger;
} hb;

int ora_use_host_blas(const char *path, const char *prefix) {
    char name[128];
    void *lib = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!lib) return -1;
#define LOAD(field, type, base)                                     \
    snprintf(name, sizeof name, "%s%s", prefix ? prefix : "", base); \
    hb.field = (type)dlsym(lib, name);                              \
    if (!hb.field) { dlclose(lib); memset(&hb, 0, sizeof hb); return -2; }
    LOAD(gemm, f_gemm, "dgemm_");
    LOAD(trmm, f_trxm, "dtrmm_");
    LOAD(trsm, f_trxm, "dtrsm_");
    LOAD(syrk, f_syrk, "dsyrk_");
    LOAD(ger, f_ger, "dger_");
#undef LOAD
    hb.lib = lib;
    return 0;
}

void ora_use_builtin_blas(void) {
    if (hb.lib) dlclose(hb.lib);
    memset(&hb, 0, sizeof hb);
}

void ora_dgemm(char ta, char tb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb,
               double beta, double *C, int ldc) {
    if (m <= 0 || n <= 0) return;
    if (hb.gemm) {
        int ldA = lda < 1 ? 1 : lda, ldB = ldb < 1 ? 1 : ldb;
        hb.gemm(&ta, &tb, &m, &n, &k, &alpha, A, &ldA, B, &ldB, &beta, C, &ldc, 1, 1);
        return;
    }
    b_dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

void ora_dtrmm(char side, char uplo, char trans, char diag, int m, int n,
               double alpha, const double *A, int lda, double *B, int ldb) {
    if (m <= 0 || n <= 0) return;
    if (hb.trmm) {
        hb.trmm(&side, &uplo, &trans, &diag, &m, &n, &alpha, A, &lda, B, &ldb, 1, 1, 1, 1);
        return;
    }
    b_dtrmm(side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}

void ora_dtrsm(char side, char uplo, char trans, char diag, int m, int n,
               double alpha, const double *A, int lda, double *B, int ldb) {
    if (m <= 0 || n <= 0) return;
    if (hb.trsm) {
        hb.trsm(&side, &uplo, &trans, &diag, &m, &n, &alpha, A, &lda, B, &ldb, 1, 1, 1, 1);
        return;
    }
    b_dtrsm(side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}

void ora_dsyrk(char uplo, char trans, int n, int k, double alpha,
               const double *A, int lda, double beta, double *C, int ldc) {
    if (n <= 0) return;
    if (hb.syrk) {
        hb.syrk(&uplo, &trans, &n, &k, &alpha, A, &lda, &beta, C, &ldc, 1, 1);
        return;
    }
    b_dsyrk(uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

void ora_dger(int m, int n, double alpha, const double *x, int incx,
              const double *y, int incy, double *A, int lda) {
    if (m <= 0 || n <= 0) return;
    if (hb.ger) {
        hb.ger(&m, &n, &alpha, x, &incx, y, &incy, A, &lda);
        return;
    }
    b_dger(m, n, alpha, x, incx, y, incy, A, lda);
}

/* ======================================================================== */
/* LAPACK subset, restated from LAPACK's published algorithms                */
/* ======================================================================== */

/* Euclidean norm with scaling (no overflow for finite input). */
static double o_dnrm2(int n, const double *x, int incx) {
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; i++) {
        double a = fabs(x[(size_t)i * incx]);
        if (a != 0.0) {
            if (scale < a) {
                ssq = 1.0 + ssq * (scale / a) * (scale / a);
                scale = a;
            } else {
                ssq += (a / scale) * (a / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}

static double o_dlapy2(double x, double y) {
    double xa = fabs(x), ya = fabs(y);
    double w = xa > ya ? xa : ya, z = xa > ya ? ya : xa;
    if (z == 0.0 || w > DBL_MAX) return w;
    return w * sqrt(1.0 + (z / w) * (z / w));
}

/* LAPACK DLARFG: H = I - tau v v^T with v(0)=1 so that H [alpha; x] = [beta; 0].
 * beta = -sign(alpha) * ||[alpha; x]||, tau = (beta-alpha)/beta,
 * v(1:) = x / (alpha - beta), with the safe-minimum rescaling loop. */
void ora_dlarfg(int n, double *alpha, double *x, int incx, double *tau) {
    if (n <= 1) { *tau = 0.0; return; }
    double xnorm = o_dnrm2(n - 1, x, incx);
    if (xnorm == 0.0) { *tau = 0.0; return; }
    double beta = -copysign(o_dlapy2(*alpha, xnorm), *alpha);
    const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
    int knt = 0;
    if (fabs(beta) < safmin) {
        const double rsafmn = 1.0 / safmin;
        do {
            knt++;
            for (int i = 0; i < n - 1; i++) x[(size_t)i * incx] *= rsafmn;
            beta *= rsafmn;
            *alpha *= rsafmn;
        } while (fabs(beta) < safmin && knt < 20);
        xnorm = o_dnrm2(n - 1, x, incx);
        beta = -copysign(o_dlapy2(*alpha, xnorm), *alpha);
    }
    *tau = (beta - *alpha) / beta;
    double sc = 1.0 / (*alpha - beta);
    for (int i = 0; i < n - 1; i++) x[(size_t)i * incx] *= sc;
    for (int j = 0; j < knt; j++) beta *= safmin;
    *alpha = beta;
}

/* LAPACK DLARFT('F','C'): T(0:i,i) = -tau_i T(0:i,0:i) V(:,0:i)^T v_i. */
void ora_dlarft_fc(int n, int k, const double *V, int ldv, const double *tau,
                   double *T, int ldt) {
    for (int i = 0; i < k; i++) {
        if (tau[i] == 0.0) {
            for (int j = 0; j <= i; j++) IDX(T, j, i, ldt) = 0.0;
            continue;
        }
        for (int j = 0; j < i; j++) {
            double s = IDX(V, i, j, ldv); /* v_i(i) = 1 */
            for (int r = i + 1; r < n; r++) s += IDX(V, r, j, ldv) * IDX(V, r, i, ldv);
            IDX(T, j, i, ldt) = -tau[i] * s;
        }
        /* T(0:i,i) = T(0:i,0:i) * T(0:i,i), upper triangular in place */
        for (int j = 0; j < i; j++) {
            double s = 0.0;
            for (int p = j; p < i; p++) s += IDX(T, j, p, ldt) * IDX(T, p, i, ldt);
            IDX(T, j, i, ldt) = s;
        }
        IDX(T, i, i, ldt) = tau[i];
    }
}

/* apply H = I - tau v v^T (v(0)=1 implicit in A(0,0) temporarily) from the
 * left to C (m x n) */
static void o_dlarf_left(int m, int n, const double *v, double tau,
                         double *C, int ldc) {
    if (tau == 0.0) return;
    for (int j = 0; j < n; j++) {
        double s = 0.0;
        for (int r = 0; r < m; r++) s += v[r] * IDX(C, r, j, ldc);
        s *= tau;
        for (int r = 0; r < m; r++) IDX(C, r, j, ldc) -= s * v[r];
    }
}

/* LAPACK DGEQR2 (unblocked Householder QR). */
void ora_dgeqr2(int m, int n, double *A, int lda, double *tau) {
    int k = m < n ? m : n;
    double *v = (double *)xmalloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    for (int i = 0; i < k; i++) {
        ora_dlarfg(m - i, &IDX(A, i, i, lda), &IDX(A, (i + 1 < m ? i + 1 : i), i, lda), 1, &tau[i]);
        if (i < n - 1) {
            v[0] = 1.0;
            for (int r = 1; r < m - i; r++) v[r] = IDX(A, i + r, i, lda);
            o_dlarf_left(m - i, n - i - 1, v, tau[i], &IDX(A, i, i + 1, lda), lda);
        }
    }
    free(v);
}

/* LAPACK DORG2R: the first n columns of H_0 ... H_{k-1}. */
void ora_dorg2r(int m, int n, int k, double *A, int lda, const double *tau) {
    double *v = (double *)xmalloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    for (int j = k; j < n; j++) {
        for (int r = 0; r < m; r++) IDX(A, r, j, lda) = 0.0;
        IDX(A, j, j, lda) = 1.0;
    }
    for (int i = k - 1; i >= 0; i--) {
        if (i < n - 1) {
            v[0] = 1.0;
            for (int r = 1; r < m - i; r++) v[r] = IDX(A, i + r, i, lda);
            o_dlarf_left(m - i, n - i - 1, v, tau[i], &IDX(A, i, i + 1, lda), lda);
        }
        for (int r = i + 1; r < m; r++) IDX(A, r, i, lda) *= -tau[i];
        IDX(A, i, i, lda) = 1.0 - tau[i];
        for (int r = 0; r < i; r++) IDX(A, r, i, lda) = 0.0;
    }
    free(v);
}

/* LAPACK DPOTRF('U') unblocked: A = U^T U.  Returns 0 or the 1-based column
 * where the matrix stopped being positive definite. */
int ora_dpotrf_u(int n, double *A, int lda) {
    for (int j = 0; j < n; j++) {
        double ajj = IDX(A, j, j, lda);
        for (int p = 0; p < j; p++) ajj -= IDX(A, p, j, lda) * IDX(A, p, j, lda);
        if (!(ajj > 0.0)) { IDX(A, j, j, lda) = ajj; return j + 1; }
        ajj = sqrt(ajj);
        IDX(A, j, j, lda) = ajj;
        for (int c = j + 1; c < n; c++) {
            double s = IDX(A, j, c, lda);
            for (int p = 0; p < j; p++) s -= IDX(A, p, j, lda) * IDX(A, p, c, lda);
            IDX(A, j, c, lda) = s / ajj;
        }
    }
    return 0;
}

void ora_dlacpy(char uplo, int m, int n, const double *A, int lda, double *B, int ldb) {
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            if ((uplo == 'U' || uplo == 'u') && i > j) continue;
            if ((uplo == 'L' || uplo == 'l') && i < j) continue;
            IDX(B, i, j, ldb) = IDX(A, i, j, lda);
        }
}

/* LAPACK DLASET: off-diagonal of the chosen part = alpha, diagonal = beta */
void ora_dlaset(char uplo, int m, int n, double alpha, double beta, double *A, int lda) {
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            if (i == j) { IDX(A, i, j, lda) = beta; continue; }
            if ((uplo == 'U' || uplo == 'u') && i > j) continue;
            if ((uplo == 'L' || uplo == 'l') && i < j) continue;
            IDX(A, i, j, lda) = alpha;
        }
}

double ora_dlange_f(int m, int n, const double *A, int lda) {
    double scale = 0.0, ssq = 1.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double a = fabs(IDX(A, i, j, lda));
            if (a != 0.0) {
                if (scale < a) { ssq = 1.0 + ssq * (scale / a) * (scale / a); scale = a; }
                else ssq += (a / scale) * (a / scale);
            }
        }
    return scale * sqrt(ssq);
}

/* ======================================================================== */
/* LILA w03 routines                                                        */
/* ======================================================================== */

/* Elmroth-Gustavson recursive panel QR; lila-c/dgeqr3.c:3-114.
 * Left half recursively, Q1^T applied to the right half with T12 as scratch
 * (steps (1)-(6), :44-70), right half recursively (:75), then
 * T12 = -T11 (V1^T V2) T22 (:95-108). */
int ora_dgeqr3(int m, int n, double *A, int lda, double *T, int ldt) {
    if (n <= 0) return 0;
    if (n == 1) {
        ora_dlarfg(m, &A[0], m > 1 ? &A[1] : &A[0], 1, &T[0]);
        return 0;
    }
    int n1 = n / 2, n2 = n - n1;
    ora_dgeqr3(m, n1, A, lda, T, ldt);

    double *W = &IDX(T, 0, n1, ldt); /* T12 doubles as the n1 x n2 scratch */
    for (int c = 0; c < n2; c++)
        for (int r = 0; r < n1; r++) IDX(W, r, c, ldt) = IDX(A, r, n1 + c, lda);
    ora_dtrmm('L', 'L', 'T', 'U', n1, n2, 1.0, A, lda, W, ldt);
    ora_dgemm('T', 'N', n1, n2, m - n1, 1.0, &IDX(A, n1, 0, lda), lda,
              &IDX(A, n1, n1, lda), lda, 1.0, W, ldt);
    ora_dtrmm('L', 'U', 'T', 'N', n1, n2, 1.0, T, ldt, W, ldt);
    ora_dgemm('N', 'N', m - n1, n2, n1, -1.0, &IDX(A, n1, 0, lda), lda, W, ldt,
              1.0, &IDX(A, n1, n1, lda), lda);
    ora_dtrmm('L', 'L', 'N', 'U', n1, n2, 1.0, A, lda, W, ldt);
    for (int c = 0; c < n2; c++)
        for (int r = 0; r < n1; r++) IDX(A, r, n1 + c, lda) -= IDX(W, r, c, ldt);

    ora_dgeqr3(m - n1, n2, &IDX(A, n1, n1, lda), lda, &IDX(T, n1, n1, ldt), ldt);

    for (int r = 0; r < n1; r++)
        for (int c = 0; c < n2; c++) IDX(T, r, n1 + c, ldt) = IDX(A, n1 + c, r, lda);
    ora_dtrmm('R', 'L', 'N', 'U', n1, n2, 1.0, &IDX(A, n1, n1, lda), lda, W, ldt);
    ora_dgemm('T', 'N', n1, n2, m - n, 1.0, &IDX(A, n, 0, lda), lda,
              &IDX(A, n, n1, lda), lda, 1.0, W, ldt);
    ora_dtrmm('L', 'U', 'N', 'N', n1, n2, -1.0, T, ldt, W, ldt);
    ora_dtrmm('R', 'U', 'N', 'N', n1, n2, 1.0, &IDX(T, n1, n1, ldt), ldt, W, ldt);
    return 0;
}

/* Per-mt-block T through DLARFT('F','C'); lila-c/lila_dlarft_w03.c:3-55.
 * First block: vb = mt - i%mt columns stored at row offset i%mt (:8-23);
 * later blocks start at row 0 of their mt-block (:29-49). */
int ora_lila_dlarft_w03(int m, int n, int i, int mt, double *A, int lda,
                        double *T, int ldt, double *tau) {
    int j = i, vb = mt - i % mt;
    if (vb > n) vb = n;
    while (vb > 0) {
        ora_dlarft_fc(m - j, vb, &IDX(A, j, j, lda), lda, tau + (j - i),
                      &IDX(T, j % mt, j, ldt), ldt);
        j += vb;
        vb = (j + mt >= i + n) ? (i + n - j) : mt;
    }
    return 0;
}

/* Join the T of the new reflectors [i, i+vb) to the wb earlier ones of the
 * same mt-block: T_ji = -T_jj (V_j^T V_i) T_ii;
 * lila-c/lila_dlarft_connect_w03.c:3-32 (wb/vb rule :8-11, steps :19-28). */
int ora_lila_dlarft_connect_w03(int m, int n, int i, int j, int mt,
                                double *A, int lda, double *T, int ldt) {
    int vb = mt - i % mt;
    if (vb > n) vb = n;
    int wb = ((j / mt) == (i / mt)) ? (i - j) : (i % mt);
    if (wb == mt) wb = 1;
    if (wb <= 0 || vb <= 0) return 0;
    const double *Aii = &IDX(A, i, i, lda);
    const double *Aij = &IDX(A, i, i - wb, lda);
    double *Tki = &IDX(T, i % mt, i, ldt);
    double *Tji = &IDX(T, i % mt - wb, i, ldt);
    const double *Tjj = &IDX(T, i % mt - wb, i - wb, ldt);
    for (int c = 0; c < vb; c++)
        for (int r = 0; r < wb; r++) IDX(Tji, r, c, ldt) = IDX(Aij, c, r, lda);
    ora_dtrmm('R', 'L', 'N', 'U', wb, vb, 1.0, Aii, lda, Tji, ldt);
    ora_dgemm('T', 'N', wb, vb, m - vb - i, 1.0, Aij + vb, lda, Aii + vb, lda, 1.0, Tji, ldt);
    ora_dtrmm('L', 'U', 'N', 'N', wb, vb, -1.0, Tjj, ldt, Tji, ldt);
    ora_dtrmm('R', 'U', 'N', 'N', wb, vb, 1.0, Tki, ldt, Tji, ldt);
    return 0;
}

/* C <- H^T C for the k reflectors starting at global i, one mt-block at a
 * time, C holding rows [i, m) of n columns.  Shared body of
 * lila-c/lila_dormqrf_w03.c:3-81 and lila-c/lila_dormqrf_z03.c:3-72:
 * W = V1^T C1 + V2^T C2; W = T^T W; C2 -= V2 W; C1 -= V1 W (:34-51), with the
 * T pointer walk of :63,:71 (block b at row b%mt, column b). */
static void o_apply_qt(int m, int n, int k, int i, int mt, const double *A,
                       int lda, const double *T, int ldt, double *C, int ldc) {
    if (k <= 0 || n <= 0) return;
    int vb = mt - i % mt;
    if (vb > k) vb = k;
    double *W = (double *)xmalloc((size_t)mt * n * sizeof(double));
    int b0 = i, done = 0;
    while (done < k) {
        const double *Vb = &IDX(A, b0, b0, lda);
        const double *Tb = &IDX(T, b0 % mt, b0, ldt);
        double *Cb = C + (b0 - i);
        int ml = m - b0;
        for (int c = 0; c < n; c++)
            for (int r = 0; r < vb; r++) IDX(W, r, c, mt) = IDX(Cb, r, c, ldc);
        ora_dtrmm('L', 'L', 'T', 'U', vb, n, 1.0, Vb, lda, W, mt);
        ora_dgemm('T', 'N', vb, n, ml - vb, 1.0, Vb + vb, lda, Cb + vb, ldc, 1.0, W, mt);
        ora_dtrmm('L', 'U', 'T', 'N', vb, n, 1.0, Tb, ldt, W, mt);
        ora_dgemm('N', 'N', ml - vb, n, vb, -1.0, Vb + vb, lda, W, mt, 1.0, Cb + vb, ldc);
        ora_dtrmm('L', 'L', 'N', 'U', vb, n, 1.0, Vb, lda, W, mt);
        for (int c = 0; c < n; c++)
            for (int r = 0; r < vb; r++) IDX(Cb, r, c, ldc) -= IDX(W, r, c, mt);
        done += vb;
        b0 += vb;
        vb = (k - done >= mt) ? mt : (k - done);
    }
    free(W);
}

int ora_lila_dormqrf_w03(int m, int n, int k, int i, int j, int mt,
                         double *A, int lda, double *T, int ldt,
                         double *work, int lwork) {
    (void)work; (void)lwork;
    o_apply_qt(m, n, k, i, mt, A, lda, T, ldt, &IDX(A, i, j, lda), lda);
    return 0;
}

int ora_lila_dormqrf_z03(int m, int n, int k, int i, int j, int mt,
                         double *A, int lda, double *T, int ldt,
                         double *B, int ldb, double *work, int lwork) {
    (void)work; (void)lwork;
    o_apply_qt(m, n, k, i, mt, A, lda, T, ldt, &IDX(B, i, j, ldb), ldb);
    return 0;
}

/* Q(:, j:j+n) <- H_i ... H_{i+k-1} [0; Q_new], walking the mt-blocks of the
 * k reflectors backwards; lila-c/lila_dormqrbz_w03.c:3-66.  Per block:
 * W = V2^T Q2 (beta 0); W = T W; Q1 = W; Q1 = -V1 Q1; Q2 -= V2 W (:25-29).
 * Positional order (A, Q, T) follows the definition, not lila-c/lila.h:34. */
int ora_lila_dormqrbz_w03(int m, int n, int k, int i, int j, int mt,
                          double *A, int lda, double *Q, int ldq,
                          double *T, int ldt, double *work, int lwork) {
    (void)work; (void)lwork;
    if (k <= 0 || n <= 0) return 0;
    /* block starts, forward: i, then the mt-aligned boundaries below i+k */
    int nblk = 0, cap = k / (mt > 0 ? mt : 1) + 2;
    int *start = (int *)xmalloc((size_t)cap * sizeof(int));
    int b0 = i, vb = mt - i % mt;
    if (vb > k) vb = k;
    while (b0 < i + k) {
        start[nblk++] = b0;
        b0 += vb;
        vb = (i + k - b0 >= mt) ? mt : (i + k - b0);
    }
    double *W = (double *)xmalloc((size_t)(mt > k ? mt : k) * n * sizeof(double));
    for (int b = nblk - 1; b >= 0; b--) {
        int s = start[b];
        int e = (b + 1 < nblk) ? start[b + 1] : (i + k);
        int w = e - s, ml = m - s;
        const double *Vb = &IDX(A, s, s, lda);
        const double *Tb = &IDX(T, s % mt, s, ldt);
        double *Qb = &IDX(Q, s, j, ldq);
        ora_dgemm('T', 'N', w, n, ml - w, 1.0, Vb + w, lda, Qb + w, ldq, 0.0, W, w);
        ora_dtrmm('L', 'U', 'N', 'N', w, n, 1.0, Tb, ldt, W, w);
        ora_dlacpy('A', w, n, W, w, Qb, ldq);
        ora_dtrmm('L', 'L', 'N', 'U', w, n, -1.0, Vb, lda, Qb, ldq);
        ora_dgemm('N', 'N', ml - w, n, w, -1.0, Vb + w, lda, W, w, 1.0, Qb + w, ldq);
    }
    free(W);
    free(start);
    return 0;
}

/* Panel = dgeqr3, then explicit panel Q = [I;0] - V (T V1^T);
 * lila-c/lila_dgeqr2_w03_3.c:3-34. */
int ora_lila_dgeqr2_w03_3(int m, int n, int i, int mt, double *A, int lda,
                          double *T, int ldt, double *Q, int ldq,
                          double *work, int lwork) {
    (void)work; (void)lwork;
    int ml = m - i;
    double *Aii = &IDX(A, i, i, lda), *Qii = &IDX(Q, i, i, ldq);
    double *Tki = &IDX(T, i % mt, i, ldt);
    ora_dgeqr3(ml, n, Aii, lda, Tki, ldt);
    double *W = (double *)xmalloc((size_t)n * n * sizeof(double));
    ora_dlacpy('A', ml, n, Aii, lda, Qii, ldq);
    ora_dlaset('U', n, n, 0.0, 1.0, Qii, ldq);
    ora_dlaset('A', n, n, 0.0, 0.0, W, n);
    ora_dlacpy('U', n, n, Tki, ldt, W, n);
    ora_dtrmm('R', 'L', 'T', 'U', n, n, 1.0, Aii, lda, W, n);
    ora_dtrmm('R', 'U', 'N', 'N', ml, n, -1.0, W, n, Qii, ldq);
    for (int c = 0; c < n; c++) IDX(Qii, c, c, ldq) += 1.0;
    free(W);
    return 0;
}

/* LAPACK panel: dgeqrf, dlarft, dorgqr; lila-c/lila_dgeqr2_w03_l.c:3-33. */
int ora_lila_dgeqr2_w03_l(int m, int n, int i, int mt, double *A, int lda,
                          double *T, int ldt, double *Q, int ldq,
                          double *work, int lwork) {
    (void)work; (void)lwork;
    int ml = m - i;
    double *tau = (double *)xmalloc((size_t)n * sizeof(double));
    double *Aii = &IDX(A, i, i, lda), *Qii = &IDX(Q, i, i, ldq);
    ora_dgeqr2(ml, n, Aii, lda, tau);
    ora_dlarft_fc(ml, n, Aii, lda, tau, &IDX(T, i % mt, i, ldt), ldt);
    ora_dlacpy('A', ml, n, Aii, lda, Qii, ldq);
    ora_dorg2r(ml, n, n, Qii, ldq, tau);
    free(tau);
    return 0;
}

/* Sign-chosen unpivoted LU of the n x n block Tq (a copy of Q1) and the
 * reconstruction of V2, signs, V1 and T = -U V1^{-T}:
 * lila-c/lila_dorgh2_3.c:8-51.  A holds R on top and Q2 below; Q is the
 * explicit Q whose columns get the sign flips (rows [0, mq)). */
static void o_reconstruct_block(int ml, int n, double *A, int lda, double *Tq,
                                int ldt, double *Q, int ldq, int mq, double *S) {
    for (int k = 0; k < n; k++) {
        double *tkk = &IDX(Tq, k, k, ldt);
        if (fabs(1.0 - *tkk) < fabs(1.0 + *tkk)) {
            for (int r = 0; r < n; r++) IDX(Tq, r, k, ldt) = -IDX(Tq, r, k, ldt);
            S[k] = -1.0;
        } else {
            S[k] = 1.0;
        }
        *tkk -= 1.0;
        for (int r = k + 1; r < n; r++) IDX(Tq, r, k, ldt) /= *tkk;
        ora_dger(n - k - 1, n - k - 1, -1.0, tkk + 1, 1, tkk + ldt, ldt, tkk + 1 + ldt, ldt);
    }
    for (int k = 0; k < n; k++)
        if (S[k] == -1.0)
            for (int r = n; r < ml; r++) IDX(A, r, k, lda) = -IDX(A, r, k, lda);
    ora_dtrsm('R', 'U', 'N', 'N', ml - n, n, 1.0, Tq, ldt, A + n, lda);
    for (int k = 0; k < n; k++)
        if (S[k] == -1.0)
            for (int c = k; c < n; c++) IDX(A, k, c, lda) = -IDX(A, k, c, lda);
    for (int k = 0; k < n; k++)
        if (S[k] == -1.0)
            for (int r = 0; r < mq; r++) IDX(Q, r, k, ldq) = -IDX(Q, r, k, ldq);
    for (int k = 0; k < n; k++)
        for (int c = 0; c < k; c++) IDX(A, k, c, lda) = IDX(Tq, k, c, ldt);
    for (int k = 0; k < n; k++)
        for (int c = 0; c < k; c++) IDX(Tq, k, c, ldt) = 0.0;
    ora_dtrsm('R', 'L', 'T', 'U', n, n, -1.0, A, lda, Tq, ldt);
}

int ora_lila_dorgh2_3(int m, int n, double *A, int lda, double *T, int ldt,
                      double *Q, int ldq, double *work, int lwork, double *S) {
    (void)work; (void)lwork;
    o_reconstruct_block(m, n, A, lda, T, ldt, Q, ldq, m, S);
    return 0;
}

/* Global-index form of dorgh2_3 in the mt layout (T block at row j%mt),
 * following storage/c2/lila_dorgh2_w03.c:3-65 with the double S of
 * lila-c/lila.h:38; Q columns j.. flipped on rows [i, m). */
int ora_lila_dorgh2(int m, int n, int i, int j, int mt, double *A, int lda,
                    double *T, int ldt, double *Q, int ldq, double *work,
                    int lwork, double *S) {
    (void)work; (void)lwork;
    o_reconstruct_block(m - j, n, &IDX(A, j, j, lda), lda, &IDX(T, j % mt, j, ldt),
                        ldt, &IDX(Q, i, j, ldq), ldq, m - i, S + j);
    return 0;
}

/* Blocked reconstruction step for block [j, j+n) given the reconstructed
 * blocks [i, j): lila-c/lila_ormhr2_w03_hr.c:3-84.  S (global index) is kept
 * apart from the workspace; in the reference S and work2 share an address
 * (lila-c/lila_dgeqrf_w03_mt_hr.c:8 vs lila-c/lila_ormhr2_w03_hr.c:10). */
int ora_lila_ormhr2_w03_hr(int m, int n, int i, int j, int l, int mt,
                           double *A, int lda, double *T, int ldt,
                           double *Q, int ldq, double *work, int lwork,
                           double *S) {
    (void)l; (void)work; (void)lwork;
    int jb = j - i;
    double *Tjj = &IDX(T, j % mt, j, ldt);
    double *Ajj = &IDX(A, j, j, lda);
    for (int k = i; k < j; k++)
        if (S[k] == -1.0)
            for (int c = 0; c < n; c++) IDX(A, k, j + c, lda) = -IDX(A, k, j + c, lda);
    double *W2 = (double *)xmalloc((size_t)(jb > 0 ? jb : 1) * n * sizeof(double));
    ora_dlacpy('A', jb, n, &IDX(Q, i, j, ldq), ldq, W2, jb > 0 ? jb : 1);
    ora_dlacpy('A', n, n, &IDX(Q, j, j, ldq), ldq, Tjj, ldt);
    ora_dlacpy('A', m - j - n, n, &IDX(Q, j + n, j, ldq), ldq, Ajj + n, lda);
    if (jb > 0) {
        ora_dtrsm('L', 'L', 'N', 'U', jb, n, 1.0, &IDX(A, i, i, lda), lda, W2, jb);
        ora_dgemm('N', 'N', n, n, jb, -1.0, &IDX(A, j, i, lda), lda, W2, jb, 1.0, Tjj, ldt);
        ora_dgemm('N', 'N', m - j - n, n, jb, -1.0, &IDX(A, j + n, i, lda), lda, W2, jb,
                  1.0, Ajj + n, lda);
    }
    free(W2);
    o_reconstruct_block(m - j, n, Ajj, lda, Tjj, ldt, &IDX(Q, i, j, ldq), ldq, m - i, S + j);
    return 0;
}

/* Blocked Householder reconstruction of A(i:m, i:i+n): Q(i:m, i:i+n)
 * orthonormal, R in the upper triangle of A's top n x n.  Per mt-block the
 * step of lila-c/lila_ormhr2_w03_hr.c (driver loop of
 * lila-c/lila_dgeqrf_w03_mt_hr.c:23-29 and storage/c2/lila_dorghr_w03.c:3-42),
 * int S of c3/lila.h:43, global index. */
int ora_lila_dorghr(int m, int n, int i, int mt, double *A, int lda,
                    double *T, int ldt, double *Q, int ldq, double *work,
                    int lwork, int *S) {
    double *Sd = (double *)xmalloc((size_t)(i + n) * sizeof(double));
    int j = i, vb = mt - i % mt;
    if (vb > n) vb = n;
    while (vb > 0) {
        ora_lila_ormhr2_w03_hr(m, vb, i, j, -1, mt, A, lda, T, ldt, Q, ldq, work, lwork, Sd);
        j += vb;
        vb = (j + mt >= i + n) ? (i + n - j) : mt;
    }
    for (int k = i; k < i + n; k++) S[k] = (int)Sd[k];
    free(Sd);
    return 0;
}

/* Cholesky-QR leaf + blocked reconstruction; lila-c/lila_dgeqrf_w03_mt_hr.c:3-33 */
int ora_lila_dgeqrf_w03_mt_hr(int m, int n, int i, int mt, double *A, int lda,
                              double *T, int ldt, double *Q, int ldq,
                              double *work, int lwork) {
    int ml = m - i;
    double *Aii = &IDX(A, i, i, lda), *Qii = &IDX(Q, i, i, ldq);
    ora_dlacpy('A', ml, n, Aii, lda, Qii, ldq);
    ora_dsyrk('U', 'T', n, ml, 1.0, Qii, ldq, 0.0, Aii, lda);
    int info = ora_dpotrf_u(n, Aii, lda);
    ora_dtrsm('R', 'U', 'N', 'N', ml, n, 1.0, Aii, lda, Qii, ldq);
    int *S = (int *)xmalloc((size_t)(i + n) * sizeof(int));
    ora_lila_dorghr(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork, S);
    free(S);
    return info;
}

/* Cholesky-QR panel + dorgh2_3; lila-c/lila_dgeqr2_w03_hr.c:3-30 */
int ora_lila_dgeqr2_w03_hr(int m, int n, int i, int mt, double *A, int lda,
                           double *T, int ldt, double *Q, int ldq,
                           double *work, int lwork) {
    int ml = m - i;
    double *Aii = &IDX(A, i, i, lda), *Qii = &IDX(Q, i, i, ldq);
    double *Tki = &IDX(T, i % mt, i, ldt);
    double *S = (double *)xmalloc((size_t)n * sizeof(double));
    ora_dlacpy('A', ml, n, Aii, lda, Qii, ldq);
    ora_dsyrk('U', 'T', n, ml, 1.0, Qii, ldq, 0.0, Aii, lda);
    int info = ora_dpotrf_u(n, Aii, lda);
    ora_dtrsm('R', 'U', 'N', 'N', ml, n, 1.0, Aii, lda, Qii, ldq);
    ora_dlacpy('A', ml - n, n, Qii + n, ldq, Aii + n, lda);
    ora_dlacpy('A', n, n, Qii, ldq, Tki, ldt);
    ora_lila_dorgh2_3(ml, n, Aii, lda, Tki, ldt, Qii, ldq, work, lwork, S);
    free(S);
    return info;
}

/* panel dispatch on lila_param[2]; lila-c/lila_dgeqr2_w03.c:8-16 */
int ora_lila_dgeqr2_w03(int *lila_param, int m, int n, int i, int mt,
                        double *A, int lda, double *T, int ldt,
                        double *Q, int ldq, double *work, int lwork) {
    switch (lila_param[2]) {
    case 0: return ora_lila_dgeqr2_w03_l(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    case 1: return ora_lila_dgeqr2_w03_3(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    case 2: return ora_lila_dgeqr2_w03_hr(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    }
    return 0;
}

/* hh leaf: mt-aligned blocks, each updated by the leaf's earlier reflectors,
 * factored, and its Q built; lila-c/lila_dgeqrf_w03_mt_hh.c:7-32 */
int ora_lila_dgeqrf_w03_mt_hh(int *lila_param, int m, int n, int i, int mt,
                              double *A, int lda, double *T, int ldt,
                              double *Q, int ldq, double *work, int lwork) {
    int vb = mt - i % mt;
    if (vb > n) vb = n;
    ora_lila_dgeqr2_w03(lila_param, m, vb, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    int j = i + vb, l = vb;
    vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
    while (vb != 0) {
        ora_lila_dormqrf_w03(m, vb, l, i, j, mt, A, lda, T, ldt, work, lwork);
        ora_lila_dgeqr2_w03(lila_param, m, vb, j, mt, A, lda, T, ldt, Q, ldq, work, lwork);
        ora_lila_dormqrbz_w03(m, vb, l, i, j, mt, A, lda, Q, ldq, T, ldt, work, lwork);
        j += vb;
        l += vb;
        vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
    }
    return 0;
}

/* LAPACK leaf: dgeqrf, lila_dlarft_w03, dorgqr; lila-c/lila_dgeqrf_w03_mt_l.c:3-25 */
int ora_lila_dgeqrf_w03_mt_l(int m, int n, int i, int mt, double *A, int lda,
                             double *T, int ldt, double *Q, int ldq,
                             double *work, int lwork) {
    (void)work; (void)lwork;
    int ml = m - i;
    double *tau = (double *)xmalloc((size_t)n * sizeof(double));
    double *Aii = &IDX(A, i, i, lda), *Qii = &IDX(Q, i, i, ldq);
    ora_dgeqr2(ml, n, Aii, lda, tau);
    ora_lila_dlarft_w03(m, n, i, mt, A, lda, T, ldt, tau);
    ora_dlacpy('A', ml, n, Aii, lda, Qii, ldq);
    ora_dorg2r(ml, n, n, Qii, ldq, tau);
    free(tau);
    return 0;
}

/* leaf dispatch on lila_param[1]; lila-c/lila_dgeqrf_w03_mt.c:6-16 */
int ora_lila_dgeqrf_w03_mt(int *lila_param, int m, int n, int i, int mt,
                           double *A, int lda, double *T, int ldt,
                           double *Q, int ldq, double *work, int lwork) {
    switch (lila_param[1]) {
    case 0: return ora_lila_dgeqrf_w03_mt_l(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    case 1: return ora_lila_dgeqrf_w03_mt_hh(lila_param, m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    case 2: return ora_lila_dgeqrf_w03_mt_hr(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    }
    return 0;
}

/* lila-c/lila_dgeqrf_w03_recursive.c:3-29 */
int ora_lila_dgeqrf_w03_recursive(int *lila_param, int m, int n, int i, int mt,
                                  double *A, int lda, double *T, int ldt,
                                  double *Q, int ldq, double *work, int lwork) {
    if (n <= lila_param[3])
        return ora_lila_dgeqrf_w03_mt(lila_param, m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    int n1 = n / 2, n2 = n - n1;
    ora_lila_dgeqrf_w03_recursive(lila_param, m, n1, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    ora_lila_dormqrf_w03(m, n2, n1, i, i + n1, mt, A, lda, T, ldt, work, lwork);
    ora_lila_dgeqrf_w03_recursive(lila_param, m, n2, i + n1, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    ora_lila_dlarft_connect_w03(m, n2, i + n1, i, mt, A, lda, T, ldt);
    ora_lila_dormqrbz_w03(m, n2, n1, i, i + n1, mt, A, lda, Q, ldq, T, ldt, work, lwork);
    return 0;
}

/* left-looking many-level blocking; lila-c/lila_dgeqrf_w03_levelx.c:3-52
 * (i_lvl passed by value, as in lila-c, not mutated as in lila-c-lila_param) */
int ora_lila_dgeqrf_w03_levelx(int *lila_param, int n_lvl, int i_lvl,
                               int *nb_lvl, int m, int n, int i, int mt,
                               double *A, int lda, double *T, int ldt,
                               double *Q, int ldq, double *work, int lwork) {
    int nb = nb_lvl[i_lvl];
    int vb = nb > n ? n : nb;
    int k = 0, j = i;
    if (i_lvl == n_lvl - 1)
        ora_lila_dgeqrf_w03_mt(lila_param, m, vb, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    else
        ora_lila_dgeqrf_w03_levelx(lila_param, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt, A, lda,
                                   T, ldt, Q, ldq, work, lwork);
    k += vb;
    j += vb;
    while (j < i + n) {
        vb = (j + nb > i + n) ? (i + n - j) : nb;
        ora_lila_dormqrf_w03(m, vb, k, i, j, mt, A, lda, T, ldt, work, lwork);
        if (i_lvl == n_lvl - 1)
            ora_lila_dgeqrf_w03_mt(lila_param, m, vb, j, mt, A, lda, T, ldt, Q, ldq, work, lwork);
        else
            ora_lila_dgeqrf_w03_levelx(lila_param, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt, A,
                                       lda, T, ldt, Q, ldq, work, lwork);
        ora_lila_dlarft_connect_w03(m, vb, j, i, mt, A, lda, T, ldt);
        ora_lila_dormqrbz_w03(m, vb, k, i, j, mt, A, lda, Q, ldq, T, ldt, work, lwork);
        k += vb;
        j += vb;
    }
    return 0;
}

/* ======================================================================== */
/* VRT-only (v03), Q-from-(V,T) (q03), append-columns, LAPACK-style and      */
/* qr2 families of lila-c/lila.h:19-22, :74-90, :95-102                     */
/* ======================================================================== */

/* LAPACK DORM2R, side 'L': C <- Q^T C (trans 'T') or Q C ('N'),
 * Q = H_0 ... H_{k-1}, H_p = I - tau_p v_p v_p^T, v_p = A(p:m, p), v_p(0) = 1 */
static void o_dorm2r(char trans, int m, int n, int k, const double *A, int lda,
                     const double *tau, double *C, int ldc) {
    double *v = (double *)xmalloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    for (int t = 0; t < k; t++) {
        int p = (trans == 'T') ? t : k - 1 - t;
        v[0] = 1.0;
        for (int r = 1; r < m - p; r++) v[r] = IDX(A, p + r, p, lda);
        o_dlarf_left(m - p, n, v, tau[p], &IDX(C, p, 0, ldc), ldc);
    }
    free(v);
}

/* dgeqr3 panel, T into the mt layout; lila-c/lila_dgeqr2_v03_3.c:3-31 (its
 * trailing T V1^T product lands in work and is never read: not restated) */
int ora_lila_dgeqr2_v03_3(int m, int n, int i, int mt, double *A, int lda,
                          double *T, int ldt, double *work, int lwork) {
    (void)work; (void)lwork;
    ora_dgeqr3(m - i, n, &IDX(A, i, i, lda), lda, &IDX(T, i % mt, i, ldt), ldt);
    return 0;
}

/* LAPACK panel: dgeqrf + dlarft; lila-c/lila_dgeqr2_v03_l.c:3-25 */
int ora_lila_dgeqr2_v03_l(int m, int n, int i, int mt, double *A, int lda,
                          double *T, int ldt, double *work, int lwork) {
    (void)work; (void)lwork;
    double *tau = (double *)xmalloc((size_t)n * sizeof(double));
    double *Aii = &IDX(A, i, i, lda);
    ora_dgeqr2(m - i, n, Aii, lda, tau);
    ora_dlarft_fc(m - i, n, Aii, lda, tau, &IDX(T, i % mt, i, ldt), ldt);
    free(tau);
    return 0;
}

/* lila-c/lila_dgeqr2_v03.c:3-19 (panel 2 is commented out there: no-op) */
int ora_lila_dgeqr2_v03(int *lila_param, int m, int n, int i, int mt, double *A,
                        int lda, double *T, int ldt, double *work, int lwork) {
    if (lila_param[2] == 0) return ora_lila_dgeqr2_v03_l(m, n, i, mt, A, lda, T, ldt, work, lwork);
    if (lila_param[2] == 1) return ora_lila_dgeqr2_v03_3(m, n, i, mt, A, lda, T, ldt, work, lwork);
    return 0;
}

/* lila-c/lila_dgeqrf_v03_mt_hh.c:3-35 */
int ora_lila_dgeqrf_v03_mt_hh(int *lila_param, int m, int n, int i, int mt,
                              double *A, int lda, double *T, int ldt,
                              double *work, int lwork) {
    int vb = mt - i % mt;
    if (vb > n) vb = n;
    ora_lila_dgeqr2_v03(lila_param, m, vb, i, mt, A, lda, T, ldt, work, lwork);
    int j = i + vb, l = vb;
    vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
    while (vb != 0) {
        ora_lila_dormqrf_w03(m, vb, l, i, j, mt, A, lda, T, ldt, work, lwork);
        ora_lila_dgeqr2_v03(lila_param, m, vb, j, mt, A, lda, T, ldt, work, lwork);
        j += vb;
        l += vb;
        vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
    }
    return 0;
}

/* lila-c/lila_dgeqrf_v03_mt_l.c:3-20 (there tau sits inside dgeqrf's own
 * workspace; here it has storage of its own) */
int ora_lila_dgeqrf_v03_mt_l(int m, int n, int i, int mt, double *A, int lda,
                             double *T, int ldt, double *work, int lwork) {
    (void)work; (void)lwork;
    double *tau = (double *)xmalloc((size_t)n * sizeof(double));
    ora_dgeqr2(m - i, n, &IDX(A, i, i, lda), lda, tau);
    ora_lila_dlarft_w03(m, n, i, mt, A, lda, T, ldt, tau);
    free(tau);
    return 0;
}

/* lila-c/lila_dgeqrf_v03_mt.c:3-19 */
int ora_lila_dgeqrf_v03_mt(int *lila_param, int m, int n, int i, int mt,
                           double *A, int lda, double *T, int ldt,
                           double *work, int lwork) {
    if (lila_param[1] == 0) return ora_lila_dgeqrf_v03_mt_l(m, n, i, mt, A, lda, T, ldt, work, lwork);
    if (lila_param[1] == 1)
        return ora_lila_dgeqrf_v03_mt_hh(lila_param, m, n, i, mt, A, lda, T, ldt, work, lwork);
    return 0;
}

/* lila-c/lila_dgeqrf_v03_recursive.c:3-29 (leaf when n < nx) */
int ora_lila_dgeqrf_v03_recursive(int *lila_param, int m, int n, int i, int mt,
                                  double *A, int lda, double *T, int ldt,
                                  double *work, int lwork) {
    if (n < lila_param[3])
        return ora_lila_dgeqrf_v03_mt(lila_param, m, n, i, mt, A, lda, T, ldt, work, lwork);
    int n1 = n / 2, n2 = n - n1;
    ora_lila_dgeqrf_v03_recursive(lila_param, m, n1, i, mt, A, lda, T, ldt, work, lwork);
    ora_lila_dormqrf_w03(m, n2, n1, i, i + n1, mt, A, lda, T, ldt, work, lwork);
    ora_lila_dgeqrf_v03_recursive(lila_param, m, n2, i + n1, mt, A, lda, T, ldt, work, lwork);
    ora_lila_dlarft_connect_w03(m, n2, i + n1, i, mt, A, lda, T, ldt);
    return 0;
}

/* lila-c/lila_dgeqrf_v03_levelx.c:3-52 */
int ora_lila_dgeqrf_v03_levelx(int *lila_param, int n_lvl, int i_lvl, int *nb_lvl,
                               int m, int n, int i, int mt, double *A, int lda,
                               double *T, int ldt, double *work, int lwork) {
    int nb = nb_lvl[i_lvl];
    int vb = nb > n ? n : nb;
    int k = 0, j = i;
    if (i_lvl == n_lvl - 1)
        ora_lila_dgeqrf_v03_mt(lila_param, m, vb, i, mt, A, lda, T, ldt, work, lwork);
    else
        ora_lila_dgeqrf_v03_levelx(lila_param, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt, A, lda, T,
                                   ldt, work, lwork);
    k += vb;
    j += vb;
    while (j < i + n) {
        vb = (j + nb > i + n) ? (i + n - j) : nb;
        ora_lila_dormqrf_w03(m, vb, k, i, j, mt, A, lda, T, ldt, work, lwork);
        if (i_lvl == n_lvl - 1)
            ora_lila_dgeqrf_v03_mt(lila_param, m, vb, j, mt, A, lda, T, ldt, work, lwork);
        else
            ora_lila_dgeqrf_v03_levelx(lila_param, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt, A, lda,
                                       T, ldt, work, lwork);
        ora_lila_dlarft_connect_w03(m, vb, j, i, mt, A, lda, T, ldt);
        k += vb;
        j += vb;
    }
    return 0;
}

/* panel Q from (V, T): Q = [I;0] - V (T V1^T); lila-c/lila_dgeqr2_q03_3.c:3-32 */
int ora_lila_dgeqr2_q03_3(int m, int n, int i, int mt, double *A, int lda,
                          double *T, int ldt, double *Q, int ldq,
                          double *work, int lwork) {
    (void)work; (void)lwork;
    int ml = m - i;
    double *Aii = &IDX(A, i, i, lda), *Qii = &IDX(Q, i, i, ldq);
    double *Tki = &IDX(T, i % mt, i, ldt);
    double *W = (double *)xmalloc((size_t)n * n * sizeof(double));
    ora_dlacpy('A', ml, n, Aii, lda, Qii, ldq);
    ora_dlaset('U', n, n, 0.0, 1.0, Qii, ldq);
    ora_dlaset('A', n, n, 0.0, 0.0, W, n);
    ora_dlacpy('U', n, n, Tki, ldt, W, n);
    ora_dtrmm('R', 'L', 'T', 'U', n, n, 1.0, Aii, lda, W, n);
    ora_dtrmm('R', 'U', 'N', 'N', ml, n, -1.0, W, n, Qii, ldq);
    for (int c = 0; c < n; c++) IDX(Qii, c, c, ldq) += 1.0;
    free(W);
    return 0;
}

/* panel Q through dorgqr with tau = diag of the T block;
 * lila-c/lila_dgeqr2_q03_l.c:3-25 */
int ora_lila_dgeqr2_q03_l(int m, int n, int i, int mt, double *A, int lda,
                          double *T, int ldt, double *Q, int ldq,
                          double *work, int lwork) {
    (void)work; (void)lwork;
    int ml = m - i;
    double *tau = (double *)xmalloc((size_t)n * sizeof(double));
    const double *Tki = &IDX(T, i % mt, i, ldt);
    for (int c = 0; c < n; c++) tau[c] = IDX(Tki, c, c, ldt);
    ora_dlacpy('A', ml, n, &IDX(A, i, i, lda), lda, &IDX(Q, i, i, ldq), ldq);
    ora_dorg2r(ml, n, n, &IDX(Q, i, i, ldq), ldq, tau);
    free(tau);
    return 0;
}

/* lila-c/lila_dgeqr2_q03.c:3-19 */
int ora_lila_dgeqr2_q03(int *lila_param, int m, int n, int i, int mt, double *A,
                        int lda, double *T, int ldt, double *Q, int ldq,
                        double *work, int lwork) {
    if (lila_param[2] == 0)
        return ora_lila_dgeqr2_q03_l(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    if (lila_param[2] == 1)
        return ora_lila_dgeqr2_q03_3(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    return 0;
}

/* lila-c/lila_dgeqrf_q03_mt_hh.c:3-32 */
int ora_lila_dgeqrf_q03_mt_hh(int *lila_param, int m, int n, int i, int mt,
                              double *A, int lda, double *T, int ldt,
                              double *Q, int ldq, double *work, int lwork) {
    int vb = mt - i % mt;
    if (vb > n) vb = n;
    ora_lila_dgeqr2_q03(lila_param, m, vb, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    int j = i + vb, l = vb;
    vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
    while (vb != 0) {
        ora_lila_dgeqr2_q03(lila_param, m, vb, j, mt, A, lda, T, ldt, Q, ldq, work, lwork);
        ora_lila_dormqrbz_w03(m, vb, l, i, j, mt, A, lda, Q, ldq, T, ldt, work, lwork);
        j += vb;
        l += vb;
        vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
    }
    return 0;
}

/* lila-c/lila_dgeqrf_q03_mt_l.c:3-23 */
int ora_lila_dgeqrf_q03_mt_l(int m, int n, int i, int mt, double *A, int lda,
                             double *T, int ldt, double *Q, int ldq,
                             double *work, int lwork) {
    (void)work; (void)lwork;
    int ml = m - i;
    double *tau = (double *)xmalloc((size_t)n * sizeof(double));
    ora_lila_dT2tau_w03(m, n, i, mt, T, ldt, tau);
    ora_dlacpy('A', ml, n, &IDX(A, i, i, lda), lda, &IDX(Q, i, i, ldq), ldq);
    ora_dorg2r(ml, n, n, &IDX(Q, i, i, ldq), ldq, tau);
    free(tau);
    return 0;
}

/* lila-c/lila_dgeqrf_q03_mt.c:3-19 */
int ora_lila_dgeqrf_q03_mt(int *lila_param, int m, int n, int i, int mt,
                           double *A, int lda, double *T, int ldt,
                           double *Q, int ldq, double *work, int lwork) {
    if (lila_param[1] == 0)
        return ora_lila_dgeqrf_q03_mt_l(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    if (lila_param[1] == 1)
        return ora_lila_dgeqrf_q03_mt_hh(lila_param, m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    return 0;
}

/* lila-c/lila_dgeqrf_q03_recursive.c:3-25 (leaf when n < nx) */
int ora_lila_dgeqrf_q03_recursive(int *lila_param, int m, int n, int i, int mt,
                                  double *A, int lda, double *T, int ldt,
                                  double *Q, int ldq, double *work, int lwork) {
    if (n < lila_param[3])
        return ora_lila_dgeqrf_q03_mt(lila_param, m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    int n1 = n / 2, n2 = n - n1;
    ora_lila_dgeqrf_q03_recursive(lila_param, m, n1, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    ora_lila_dgeqrf_q03_recursive(lila_param, m, n2, i + n1, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    ora_lila_dormqrbz_w03(m, n2, n1, i, i + n1, mt, A, lda, Q, ldq, T, ldt, work, lwork);
    return 0;
}

/* lila-c/lila_dgeqrf_q03_levelx.c:3-48 */
int ora_lila_dgeqrf_q03_levelx(int *lila_param, int n_lvl, int i_lvl, int *nb_lvl,
                               int m, int n, int i, int mt, double *A, int lda,
                               double *T, int ldt, double *Q, int ldq,
                               double *work, int lwork) {
    int nb = nb_lvl[i_lvl];
    int vb = nb > n ? n : nb;
    int k = 0, j = i;
    if (i_lvl == n_lvl - 1)
        ora_lila_dgeqrf_q03_mt(lila_param, m, vb, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    else
        ora_lila_dgeqrf_q03_levelx(lila_param, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt, A, lda, T,
                                   ldt, Q, ldq, work, lwork);
    k += vb;
    j += vb;
    while (j < i + n) {
        vb = (j + nb > i + n) ? (i + n - j) : nb;
        if (i_lvl == n_lvl - 1)
            ora_lila_dgeqrf_q03_mt(lila_param, m, vb, j, mt, A, lda, T, ldt, Q, ldq, work, lwork);
        else
            ora_lila_dgeqrf_q03_levelx(lila_param, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt, A, lda,
                                       T, ldt, Q, ldq, work, lwork);
        ora_lila_dormqrbz_w03(m, vb, k, i, j, mt, A, lda, Q, ldq, T, ldt, work, lwork);
        k += vb;
        j += vb;
    }
    return 0;
}

/* QR of k existing columns extended by n new ones (Block-Arnoldi use):
 * lila-c/lila_dgeqrf_w03_appendcols.c:3-14 */
int ora_lila_dgeqrf_w03_appendcols(int *lila_param, int m, int k, int n, int mt,
                                   double *A, int lda, double *Q, int ldq,
                                   double *T, int ldt, double *work, int lwork) {
    if (k != 0) ora_lila_dormqrf_w03(m, n, k, 0, k, mt, A, lda, T, ldt, work, lwork);
    ora_lila_dgeqrf_w03_recursive(lila_param, m, n, k, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    if (k != 0) ora_lila_dlarft_connect_w03(m, n, k, 0, mt, A, lda, T, ldt);
    if (k != 0) ora_lila_dormqrbz_w03(m, n, k, 0, k, mt, A, lda, Q, ldq, T, ldt, work, lwork);
    return 0;
}

/* the same through LAPACK conventions, tau stored in T (a vector of k + n):
 * lila-c/lila_dgeqrf_LAPACK_appendcols.c:3-30.  The reference passes ldq as
 * the leading dimension of A's new columns in its first dormqr (:20); the
 * two agree when lda == ldq, and lda is used here. */
int ora_lila_dgeqrf_LAPACK_appendcols(int m, int k, int n, int mt, double *A,
                                      int lda, double *Q, int ldq, double *T,
                                      int ldt, double *work, int lwork) {
    (void)mt; (void)ldt; (void)work; (void)lwork;
    double *tau = T;
    int ml = m - k;
    double *Aji = &IDX(A, 0, k, lda), *Qji = &IDX(Q, 0, k, ldq);
    if (k != 0) o_dorm2r('T', m, n, k, A, lda, tau, Aji, lda);
    ora_dgeqr2(ml, n, &IDX(A, k, k, lda), lda, tau + k);
    ora_dlacpy('A', ml, n, &IDX(A, k, k, lda), lda, &IDX(Q, k, k, ldq), ldq);
    ora_dorg2r(ml, n, n, &IDX(Q, k, k, ldq), ldq, tau + k);
    if (k != 0) ora_dlaset('A', k, n, 0.0, 0.0, Qji, ldq);
    if (k != 0) o_dorm2r('N', m, n, k, A, lda, tau, Qji, ldq);
    return 0;
}

/* recursive QR on LAPACK kernels, tau stored in T (T + i holds the tau of
 * the columns from i): lila-c/lila_dgeqrf_ker_recursive.c:3-44.  The
 * reference's dormqr on A's right half passes ldq as its leading dimension
 * (:31); lda is used here (the same when lda == ldq). */
int ora_lila_dgeqrf_ker_recursive(int *lila_param, int m, int n, int i, int mt,
                                  double *A, int lda, double *T, int ldt,
                                  double *Q, int ldq, double *work, int lwork) {
    (void)ldt;
    double *tau = T + i;
    double *Aii = &IDX(A, i, i, lda), *Qii = &IDX(Q, i, i, ldq);
    int ml = m - i;
    if (n <= lila_param[3]) {
        ora_dgeqr2(ml, n, Aii, lda, tau);
        ora_dlacpy('A', ml, n, Aii, lda, Qii, ldq);
        ora_dorg2r(ml, n, n, Qii, ldq, tau);
        return 0;
    }
    int n1 = n / 2, n2 = n - n1;
    ora_lila_dgeqrf_ker_recursive(lila_param, m, n1, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    o_dorm2r('T', ml, n2, n1, Aii, lda, tau, Aii + (size_t)n1 * lda, lda);
    ora_lila_dgeqrf_ker_recursive(lila_param, m, n2, i + n1, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    ora_dlaset('A', n1, n2, 0.0, 0.0, Qii + (size_t)n1 * ldq, ldq);
    o_dorm2r('N', ml, n2, n1, Aii, lda, tau, Qii + (size_t)n1 * ldq, ldq);
    return 0;
}

/* tau_p = 2 / (1 + ||v_p(1:)||^2) from the stored reflectors, then dormqr:
 * lila-c/lila_dormqrf_w00.c:3-29 (C = A(i:m, j:j+n)) and
 * lila-c/lila_dormqrf_z00.c:3-31 (C = B(i:m, j:j+n)) */
static void o_dormqrf_tau_from_v(int m, int n, int k, int i, const double *A, int lda,
                                 double *C, int ldc) {
    int ml = m - i;
    double *tau = (double *)xmalloc((size_t)(k > 0 ? k : 1) * sizeof(double));
    for (int p = 0; p < k; p++) {
        double s = 1.0;
        for (int r = i + p + 1; r < m; r++) s += IDX(A, r, i + p, lda) * IDX(A, r, i + p, lda);
        tau[p] = 2.0 / s;
    }
    o_dorm2r('T', ml, n, k, &IDX(A, i, i, lda), lda, tau, C, ldc);
    free(tau);
}

int ora_lila_dormqrf_w00(int m, int n, int k, int i, int j, int mt, double *A, int lda,
                         double *T, int ldt, double *work, int lwork) {
    (void)mt; (void)T; (void)ldt; (void)work; (void)lwork;
    o_dormqrf_tau_from_v(m, n, k, i, A, lda, &IDX(A, i, j, lda), lda);
    return 0;
}

int ora_lila_dormqrf_z00(int m, int n, int k, int i, int j, int mt, double *A, int lda,
                         double *B, int ldb, double *T, int ldt, double *work, int lwork) {
    (void)mt; (void)T; (void)ldt; (void)work; (void)lwork;
    o_dormqrf_tau_from_v(m, n, k, i, A, lda, &IDX(B, i, j, ldb), ldb);
    return 0;
}

/* one block of k reflectors with the full-layout T at T(i, i):
 * lila-c/lila_dormqrf_z02.c:3-42 */
int ora_lila_dormqrf_z02(int m, int n, int k, int i, int j, int mt, double *A, int lda,
                         double *T, int ldt, double *B, int ldb, double *work, int lwork) {
    (void)mt; (void)work; (void)lwork;
    const double *Aii = &IDX(A, i, i, lda), *Tii = &IDX(T, i, i, ldt);
    double *Bij = &IDX(B, i, j, ldb);
    double *W = (double *)xmalloc((size_t)(k > 0 ? k : 1) * n * sizeof(double));
    for (int c = 0; c < n; c++)
        for (int r = 0; r < k; r++) IDX(W, r, c, k) = IDX(Bij, r, c, ldb);
    ora_dtrmm('L', 'L', 'T', 'U', k, n, 1.0, Aii, lda, W, k);
    ora_dgemm('T', 'N', k, n, m - k - i, 1.0, Aii + k, lda, Bij + k, ldb, 1.0, W, k);
    ora_dtrmm('L', 'U', 'T', 'N', k, n, 1.0, Tii, ldt, W, k);
    ora_dgemm('N', 'N', m - k - i, n, k, -1.0, Aii + k, lda, W, k, 1.0, Bij + k, ldb);
    ora_dtrmm('L', 'L', 'N', 'U', k, n, 1.0, Aii, lda, W, k);
    for (int c = 0; c < n; c++)
        for (int r = 0; r < k; r++) IDX(Bij, r, c, ldb) -= IDX(W, r, c, k);
    free(W);
    return 0;
}

/* panel by lila_param[4] (0 VRTQ, 1 VRT, 2 Q) and lila_param[2];
 * lila-c/lila_dgeqrf_qr2.c:3-45 (the hr forms of 1 and 2 are commented out
 * there: no-ops) */
int ora_lila_dgeqrf_qr2(int *lila_param, int m, int n, int i, int mt, double *A, int lda,
                        double *T, int ldt, double *Q, int ldq, double *work, int lwork) {
    int panel = lila_param[2];
    if (lila_param[4] == 0) {
        if (panel == 0) return ora_lila_dgeqr2_w03_l(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
        if (panel == 1) return ora_lila_dgeqr2_w03_3(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
        if (panel == 2) return ora_lila_dgeqr2_w03_hr(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    } else if (lila_param[4] == 1) {
        if (panel == 0) return ora_lila_dgeqr2_v03_l(m, n, i, mt, A, lda, T, ldt, work, lwork);
        if (panel == 1) return ora_lila_dgeqr2_v03_3(m, n, i, mt, A, lda, T, ldt, work, lwork);
    } else if (lila_param[4] == 2) {
        if (panel == 0) return ora_lila_dgeqr2_q03_l(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
        if (panel == 1) return ora_lila_dgeqr2_q03_3(m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    }
    return 0;
}

/* lila-c/lila_dgeqrf_qr2_recursive.c:3-24 */
int ora_lila_dgeqrf_qr2_recursive(int *lila_param, int m, int n, int i, int mt, double *A,
                                  int lda, double *T, int ldt, double *Q, int ldq,
                                  double *work, int lwork) {
    if (n <= lila_param[3])
        return ora_lila_dgeqrf_qr2(lila_param, m, n, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
    int n1 = n / 2, n2 = n - n1;
    return ora_lila_dgeqrf_qr2_recursive2(lila_param, m, n, i, mt, n1, n2, A, lda, T, ldt, Q, ldq,
                                          work, lwork);
}

/* the three step lists of lila-c/lila_dgeqrf_qr2_recursive2.c:3-34 */
int ora_lila_dgeqrf_qr2_recursive2(int *lila_param, int m, int n, int i, int mt, int nb1, int nb2,
                                   double *A, int lda, double *T, int ldt, double *Q, int ldq,
                                   double *work, int lwork) {
    (void)n;
    int method = lila_param[4];
    if (method == 0 || method == 1) {
        ora_lila_dgeqrf_qr2_recursive(lila_param, m, nb1, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
        ora_lila_dormqrf_w03(m, nb2, nb1, i, i + nb1, mt, A, lda, T, ldt, work, lwork);
        ora_lila_dgeqrf_qr2_recursive(lila_param, m, nb2, i + nb1, mt, A, lda, T, ldt, Q, ldq, work,
                                      lwork);
        ora_lila_dlarft_connect_w03(m, nb2, i + nb1, i, mt, A, lda, T, ldt);
        if (method == 0)
            ora_lila_dormqrbz_w03(m, nb2, nb1, i, i + nb1, mt, A, lda, Q, ldq, T, ldt, work, lwork);
    } else if (method == 2) {
        ora_lila_dgeqrf_qr2_recursive(lila_param, m, nb1, i, mt, A, lda, T, ldt, Q, ldq, work, lwork);
        ora_lila_dgeqrf_qr2_recursive(lila_param, m, nb2, i + nb1, mt, A, lda, T, ldt, Q, ldq, work,
                                      lwork);
        ora_lila_dormqrbz_w03(m, nb2, nb1, i, i + nb1, mt, A, lda, Q, ldq, T, ldt, work, lwork);
    }
    return 0;
}

/* In-place sign-chosen unpivoted LU of A(i:m, i:i+n) (holding a copy of the
 * orthonormal Q): per column, if |1 - a_kk| < |1 + a_kk| negate row k of R
 * (held in work, n x n, ld n) and column k of Q(i:m, i:i+n), otherwise negate
 * column k of A; then a_kk += 1, scale below, rank-1 update of the
 * remaining columns; c3/lila_dgetrf_b.c:3-34. */
int ora_lila_dgetrf_b(int m, int n, int i, int mt, double *A, int lda,
                      double *Q, int ldq, double *work, int lwork) {
    (void)mt; (void)lwork;
    int ml = m - i;
    double *Aii = &IDX(A, i, i, lda), *Qii = &IDX(Q, i, i, ldq), *U = work;
    for (int k = 0; k < n; k++) {
        double akk = IDX(Aii, k, k, lda);
        if (fabs(1.0 - akk) < fabs(1.0 + akk)) {
            for (int c = k; c < n; c++) IDX(U, k, c, n) = -IDX(U, k, c, n);
            for (int r = 0; r < ml; r++) IDX(Qii, r, k, ldq) = -IDX(Qii, r, k, ldq);
        } else {
            for (int r = 0; r < ml; r++) IDX(Aii, r, k, lda) = -IDX(Aii, r, k, lda);
        }
        IDX(Aii, k, k, lda) += 1.0;
        double piv = IDX(Aii, k, k, lda);
        for (int r = k + 1; r < ml; r++) IDX(Aii, r, k, lda) /= piv;
        ora_dgemm('N', 'N', ml - k - 1, n - k - 1, 1, -1.0, &IDX(Aii, k + 1, k, lda), lda,
                  &IDX(Aii, k, k + 1, lda), lda, 1.0, &IDX(Aii, k + 1, k + 1, lda), lda);
    }
    return 0;
}

/* tau from the diagonal of the mt-blocked T; lila-c/lila_dT2tau_w03.c:3-35 */
int ora_lila_dT2tau_w03(int m, int n, int i, int mt, double *T, int ldt, double *tau) {
    (void)m;
    for (int c = 0; c < n; c++) {
        int g = i + c;
        tau[c] = IDX(T, g % mt, g, ldt);
    }
    return 0;
}

/* tau = 2 / (1 + ||v_2||^2) per reflector; lila-c/lila_dV2tau_w03.c:12.
 * The reflector of global column g has v_2 = A(g+1:m, g).  (The reference's
 * later-block loop keeps ml at its first-block value, :24-31, so it reads
 * past the column; the restatement uses the column length m-g-1.) */
int ora_lila_dV2tau_w03(int m, int n, int i, int mt, double *A, int lda, double *tau) {
    (void)mt;
    for (int c = 0; c < n; c++) {
        int g = i + c;
        double s = 0.0;
        for (int r = g + 1; r < m; r++) s += IDX(A, r, g, lda) * IDX(A, r, g, lda);
        tau[c] = 2.0 / (1.0 + s);
    }
    return 0;
}

/* ---- workspace queries (lila-c/lila_query_*.c) -------------------------- */

/* lila-c/lila_query_dormqrf_w03.c */
int ora_lila_query_dormqrf_w03(int m, int n, int k, int i, int j, int mt,
                               double *A, int lda, double *T, int ldt,
                               double *work, int lwork) {
    (void)j; (void)A; (void)lda; (void)T; (void)ldt; (void)work; (void)lwork;
    int vb = mt - i % mt;
    if (vb > k) vb = k;
    int w = 0, ml = m - i, jj = 1;
    if (w < n * mt) w = n * mt;
    for (;;) {
        if (w < (ml - vb) * n) w = (ml - vb) * n;
        if (jj + vb - 1 >= k) break;
        ml -= vb;
        jj += vb;
        vb = (jj + mt - 1 <= k) ? mt : (k - jj + 1);
    }
    return w;
}

/* lila-c/lila_query_dormqrbz_w03.c */
int ora_lila_query_dormqrbz_w03(int m, int n, int k, int i, int j, int mt,
                                double *A, int lda, double *Q, int ldq,
                                double *T, int ldt, double *work, int lwork) {
    (void)j; (void)A; (void)lda; (void)Q; (void)ldq; (void)T; (void)ldt; (void)work; (void)lwork;
    int vb = (i + k) % mt;
    if (vb == 0) vb = mt;
    if (vb > k) vb = k;
    int w = n * k, ml = m - (i + k - vb), jj = 1;
    while (jj + vb - 1 != k) {
        if (jj + vb - 1 > k) break;
        if (w < n * vb) w = n * vb;
        if (w < (ml - vb) * k) w = (ml - vb) * k;
        jj += vb;
        vb = (jj + mt - 1 <= k) ? mt : (k - jj + 1);
        ml += vb;
    }
    return w;
}

static int q_dgeqr2(int *p, int n) {
    if (p[2] == 0) return n * n; /* lila-c/lila_query_dgeqr2_w03_l.c */
    if (p[2] == 1) return n * n; /* lila-c/lila_query_dgeqr2_w03_3.c */
    if (p[2] == 2) return n;     /* lila-c/lila_query_dgeqr2_w03_hr.c */
    return 0;
}

/* lila-c/lila_query_dgeqrf_w03_mt.c and _mt_hh.c, _mt_l.c, _mt_hr.c */
int ora_lila_query_dgeqrf_w03_mt(int *lila_param, int m, int n, int i, int mt,
                                 double *A, int lda, double *T, int ldt,
                                 double *Q, int ldq, double *work, int lwork) {
    int w = lwork;
    if (lila_param[1] == 0) {
        if (w < n * n) w = n * n;
    } else if (lila_param[1] == 1) {
        int vb = mt - i % mt;
        if (vb > n) vb = n;
        int q = q_dgeqr2(lila_param, vb);
        if (w < q) w = q;
        int j = i + vb, l = vb;
        vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
        while (vb != 0) {
            q = ora_lila_query_dormqrf_w03(m, vb, l, i, j, mt, A, lda, T, ldt, work, w);
            if (w < q) w = q;
            q = q_dgeqr2(lila_param, vb);
            if (w < q) w = q;
            q = ora_lila_query_dormqrbz_w03(m, vb, l, i, j, mt, A, lda, Q, ldq, T, ldt, work, w);
            if (w < q) w = q;
            j += vb;
            l += vb;
            vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
        }
    } else if (lila_param[1] == 2) {
        int j = i, vb = mt - i % mt;
        if (vb > n) vb = n;
        while (vb != 0) {
            int q = vb * (j - i) + vb; /* lila-c/lila_query_ormhr2_w03_hr.c */
            if (w < q) w = q;
            j += vb;
            vb = (j + mt >= i + n) ? (n - (j - i)) : mt;
        }
    }
    return w;
}

/* lila-c/lila_query_dgeqrf_w03_recursive.c */
int ora_lila_query_dgeqrf_w03_recursive(int *lila_param, int m, int n, int i,
                                        int mt, double *A, int lda, double *T,
                                        int ldt, double *Q, int ldq,
                                        double *work, int lwork) {
    int w = lwork, q;
    if (n <= lila_param[3]) {
        q = ora_lila_query_dgeqrf_w03_mt(lila_param, m, n, i, mt, A, lda, T, ldt, Q, ldq, work, w);
        if (w < q) w = q;
    } else {
        int n1 = n / 2, n2 = n - n1;
        q = ora_lila_query_dgeqrf_w03_recursive(lila_param, m, n1, i, mt, A, lda, T, ldt, Q, ldq, work, w);
        if (w < q) w = q;
        q = ora_lila_query_dormqrf_w03(m, n2, n1, i, i + n1, mt, A, lda, T, ldt, work, w);
        if (w < q) w = q;
        q = ora_lila_query_dgeqrf_w03_recursive(lila_param, m, n2, i + n1, mt, A, lda, T, ldt, Q, ldq, work, w);
        if (w < q) w = q;
        q = ora_lila_query_dormqrbz_w03(m, n2, n1, i, i + n1, mt, A, lda, Q, ldq, T, ldt, work, w);
        if (w < q) w = q;
    }
    return w;
}

/* lila-c/lila_query_dgeqrf_w03_levelx.c (its leaf query uses the full n) */
int ora_lila_query_dgeqrf_w03_levelx(int *lila_param, int n_lvl, int i_lvl,
                                     int *nb_lvl, int m, int n, int i, int mt,
                                     double *A, int lda, double *T, int ldt,
                                     double *Q, int ldq, double *work,
                                     int lwork) {
    (void)lwork;
    int w = 0, q;
    int nb = nb_lvl[i_lvl];
    int vb = nb > n ? n : nb;
    int k = 0, j = i;
    if (i_lvl == n_lvl - 1)
        q = ora_lila_query_dgeqrf_w03_mt(lila_param, m, n, i, mt, A, lda, T, ldt, Q, ldq, work, w);
    else
        q = ora_lila_query_dgeqrf_w03_levelx(lila_param, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt,
                                             A, lda, T, ldt, Q, ldq, work, w);
    if (q > w) w = q;
    k += vb;
    j += vb;
    while (j < i + n) {
        vb = (j + nb > i + n) ? (i + n - j) : nb;
        q = ora_lila_query_dormqrf_w03(m, vb, k, i, j, mt, A, lda, T, ldt, work, w);
        if (q > w) w = q;
        if (i_lvl == n_lvl - 1)
            q = ora_lila_query_dgeqrf_w03_mt(lila_param, m, n, i, mt, A, lda, T, ldt, Q, ldq, work, w);
        else
            q = ora_lila_query_dgeqrf_w03_levelx(lila_param, n_lvl, i_lvl + 1, nb_lvl, m, vb, j, mt,
                                                 A, lda, T, ldt, Q, ldq, work, w);
        if (q > w) w = q;
        q = ora_lila_query_dormqrbz_w03(m, vb, k, i, j, mt, A, lda, Q, ldq, T, ldt, work, w);
        if (q > w) w = q;
        k += vb;
        j += vb;
    }
    return w;
}

/* ---- least squares (least-squares/main_ls_qr3.c:94-119) ---------------- */
int ora_ls_qr3(int m, int n, int k, double *A, int lda, double *T, int ldt,
               double *B, int ldb, double *work) {
    ora_dgeqr3(m, n, A, lda, T, ldt);
    ora_dlacpy('A', n, k, B, ldb, work, n);
    ora_dtrmm('L', 'L', 'T', 'U', n, k, 1.0, A, lda, work, n);
    ora_dgemm('T', 'N', n, k, m - n, 1.0, A + n, lda, B + n, ldb, 1.0, work, n);
    ora_dtrmm('L', 'U', 'T', 'N', n, k, 1.0, T, ldt, work, n);
    ora_dgemm('N', 'N', m - n, k, n, -1.0, A + n, lda, work, n, 1.0, B + n, ldb);
    ora_dtrmm('L', 'L', 'N', 'U', n, k, 1.0, A, lda, work, n);
    for (int c = 0; c < k; c++)
        for (int r = 0; r < n; r++) IDX(B, r, c, ldb) -= IDX(work, r, c, n);
    ora_dtrsm('L', 'U', 'N', 'N', n, k, 1.0, A, lda, B, ldb);
    return 0;
}

/* LAPACK DGELS('N') for m >= n, the per-problem host baseline of the batched
 * least-squares workload (least-squares/main_lapack_ref.c:100 calls dgels):
 * dgeqr2 (A <- V below the diagonal, R on and above), B <- Q^T B (dorm2r),
 * then R X = B(0:n, :) by back substitution (the dtrsm of
 * least-squares/main_ls_qr3.c:117).  tau (n doubles) is returned too. */
int ora_dgels_hh(int m, int n, int k, double *A, int lda, double *B, int ldb, double *tau) {
    if (m < n || n < 0 || k < 0) return -1;
    ora_dgeqr2(m, n, A, lda, tau);
    if (k > 0) {
        o_dorm2r('T', m, k, n, A, lda, tau, B, ldb);
        ora_dtrsm('L', 'U', 'N', 'N', n, k, 1.0, A, lda, B, ldb);
    }
    return 0;
}

/* ---- the recursive QR library of qr3/src (qr2.h) ------------------------ */

/* qr3/src/qr2_dgeqr3_R.c:3-60 (R apart from A; the W workspace is T12 when
 * R aliases A, else R12) */
int ora_qr2_dgeqr3_R(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr) {
    if (n == 1) {
        R[0] = A[0];
        ora_dlarfg(m, R, A + 1, 1, T);
        return 0;
    }
    int n1 = n / 2, n2 = n - n1, ldw;
    double *A11 = A, *A12 = A + (size_t)n1 * lda, *A21 = A + n1, *A22 = A + n1 + (size_t)n1 * lda;
    double *T11 = T, *T12 = T + (size_t)n1 * ldt, *T22 = T + n1 + (size_t)n1 * ldt;
    double *R11 = R, *R12 = R + (size_t)n1 * ldr, *R22 = R + n1 + (size_t)n1 * ldr, *W;
    ora_qr2_dgeqr3_R(m, n1, A11, lda, T11, ldt, R11, ldr);
    if (A11 == R11) { W = T12; ldw = ldt; } else { W = R12; ldw = ldr; }
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) W[i + (size_t)j * ldw] = A12[(size_t)j * lda + i];
    ora_dtrmm('L', 'L', 'T', 'U', n1, n2, 1.0, A11, lda, W, ldw);
    ora_dgemm('T', 'N', n1, n2, m - n1, 1.0, A21, lda, A22, lda, 1.0, W, ldw);
    ora_dtrmm('L', 'U', 'T', 'N', n1, n2, 1.0, T11, ldt, W, ldw);
    ora_dgemm('N', 'N', m - n1, n2, n1, -1.0, A21, lda, W, ldw, 1.0, A22, lda);
    ora_dtrmm('L', 'L', 'N', 'U', n1, n2, 1.0, A11, lda, W, ldw);
    for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) R12[(size_t)j * ldr + i] = A12[(size_t)j * lda + i] - W[(size_t)j * ldw + i];
    ora_qr2_dgeqr3_R(m - n1, n2, A22, lda, T22, ldt, R22, ldr);
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) T12[i + (size_t)j * ldt] = A21[j + (size_t)i * lda];
    ora_dtrmm('R', 'L', 'N', 'U', n1, n2, 1.0, A22, lda, T12, ldt);
    ora_dgemm('T', 'N', n1, n2, m - n, 1.0, A21 + n2, lda, A22 + n2, lda, 1.0, T12, ldt);
    ora_dtrmm('L', 'U', 'N', 'N', n1, n2, -1.0, T11, ldt, T12, ldt);
    ora_dtrmm('R', 'U', 'N', 'N', n1, n2, 1.0, T22, ldt, T12, ldt);
    return 0;
}

/* qr3/src/qr2_dlarft3.c:3-44 */
int ora_qr2_dlarft3(int m, int n, double *A, int lda, double *T, int ldt, double *tau) {
    if (n <= 1) {
        T[0] = tau[0];
        return 0;
    }
    int n1 = n / 2, n2 = n - n1;
    double *A21 = A + n1, *A31 = A + n, *A22 = A + n1 + (size_t)n1 * lda, *A32 = A + n + (size_t)n1 * lda;
    double *T12 = T + (size_t)n1 * ldt, *T22 = T + n1 + (size_t)n1 * ldt;
    ora_qr2_dlarft3(m, n1, A, lda, T, ldt, tau);
    ora_qr2_dlarft3(m - n1, n2, A22, lda, T22, ldt, tau + n1);
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) T12[i + (size_t)j * ldt] = A21[j + (size_t)i * lda];
    ora_dtrmm('R', 'L', 'N', 'U', n1, n2, 1.0, A22, lda, T12, ldt);
    ora_dgemm('T', 'N', n1, n2, m - n, 1.0, A31, lda, A32, lda, 1.0, T12, ldt);
    ora_dtrmm('L', 'U', 'N', 'N', n1, n2, -1.0, T, ldt, T12, ldt);
    ora_dtrmm('R', 'U', 'N', 'N', n1, n2, 1.0, T22, ldt, T12, ldt);
    return 0;
}

/* qr3/src/qr2_dorgqr3.c:3-58 */
int ora_qr2_dorgqr3(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau) {
    if (n <= 1) {
        for (int i = 1; i < m; i++) Q[i] *= -tau[0];
        Q[0] = 1.0 - tau[0];
        return 0;
    }
    int n1 = n / 2, n2 = n - n1;
    double *Q21 = Q + n1, *Q12 = Q + (size_t)n1 * ldq, *Q22 = Q + n1 + (size_t)n1 * ldq;
    double *T22 = T + n1 + (size_t)n1 * ldt;
    ora_qr2_dorgqr3(m - n1, n2, Q22, ldq, T22, ldt, tau + n1);
    ora_dgemm('T', 'N', n1, n2, m - n1, 1.0, Q21, ldq, Q22, ldq, 0.0, Q12, ldq);
    ora_dtrmm('L', 'U', 'N', 'N', n1, n2, 1.0, T, ldt, Q12, ldq);
    ora_dgemm('N', 'N', m - n1, n2, n1, -1.0, Q21, ldq, Q12, ldq, 1.0, Q22, ldq);
    ora_dtrmm('L', 'L', 'N', 'U', n1, n2, -1.0, Q, ldq, Q12, ldq);
    ora_qr2_dorgqr3(m, n1, Q, ldq, T, ldt, tau);
    return 0;
}

/* The dgeqr3_R family with the form flags: ut = 1 carries S = T^-1
 * (qr3/src/qr2_dgeqr3_R_UT.c:3-67: leaf S = 1/tau, the update solves with
 * S11^T where the T form multiplies by T11^T, and the off-diagonal block
 * keeps V1^T V2); isw = 1 skips the off-diagonal block, the left half
 * running the non-ISW routine of the same form (qr2_dgeqr3_R_ISW.c:3-63,
 * qr2_dgeqr3_R_UT_ISW.c:3-64).  form(0, 0) is ora_qr2_dgeqr3_R. */
static void o_geqr3_R_form(int ut, int isw, int m, int n, double *A, int lda, double *T, int ldt,
                           double *R, int ldr) {
    if (n == 1) {
        R[0] = A[0];
        ora_dlarfg(m, R, A + 1, 1, T);
        if (ut) T[0] = 1.0 / T[0];
        return;
    }
    const int n1 = n / 2, n2 = n - n1;
    double *A12 = A + (size_t)n1 * lda, *A21 = A + n1, *A22 = A + n1 + (size_t)n1 * lda;
    double *T12 = T + (size_t)n1 * ldt, *T22 = T + n1 + (size_t)n1 * ldt;
    double *R12 = R + (size_t)n1 * ldr, *R22 = R + n1 + (size_t)n1 * ldr;
    o_geqr3_R_form(ut, 0, m, n1, A, lda, T, ldt, R, ldr);
    double *W = (A == R) ? T12 : R12;
    const int ldw = (A == R) ? ldt : ldr;
    ora_dlacpy('A', n1, n2, A12, lda, W, ldw);
    ora_dtrmm('L', 'L', 'T', 'U', n1, n2, 1.0, A, lda, W, ldw);
    ora_dgemm('T', 'N', n1, n2, m - n1, 1.0, A21, lda, A22, lda, 1.0, W, ldw);
    if (ut) ora_dtrsm('L', 'U', 'T', 'N', n1, n2, 1.0, T, ldt, W, ldw);
    else ora_dtrmm('L', 'U', 'T', 'N', n1, n2, 1.0, T, ldt, W, ldw);
    ora_dgemm('N', 'N', m - n1, n2, n1, -1.0, A21, lda, W, ldw, 1.0, A22, lda);
    ora_dtrmm('L', 'L', 'N', 'U', n1, n2, 1.0, A, lda, W, ldw);
    for (int c = 0; c < n2; c++)
        for (int r = 0; r < n1; r++) IDX(R12, r, c, ldr) = IDX(A12, r, c, lda) - IDX(W, r, c, ldw);
    o_geqr3_R_form(ut, isw, m - n1, n2, A22, lda, T22, ldt, R22, ldr);
    if (isw) return;
    for (int r = 0; r < n1; r++)
        for (int c = 0; c < n2; c++) IDX(T12, r, c, ldt) = IDX(A21, c, r, lda);
    ora_dtrmm('R', 'L', 'N', 'U', n1, n2, 1.0, A22, lda, T12, ldt);
    ora_dgemm('T', 'N', n1, n2, m - n, 1.0, A21 + n2, lda, A22 + n2, lda, 1.0, T12, ldt);
    if (!ut) {
        ora_dtrmm('L', 'U', 'N', 'N', n1, n2, -1.0, T, ldt, T12, ldt);
        ora_dtrmm('R', 'U', 'N', 'N', n1, n2, 1.0, T22, ldt, T12, ldt);
    }
}

int ora_qr2_dgeqr3_R_UT(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr) {
    o_geqr3_R_form(1, 0, m, n, A, lda, T, ldt, R, ldr);
    return 0;
}
int ora_qr2_dgeqr3_R_ISW(int m, int n, double *A, int lda, double *T, int ldt, double *R, int ldr) {
    o_geqr3_R_form(0, 1, m, n, A, lda, T, ldt, R, ldr);
    return 0;
}
int ora_qr2_dgeqr3_R_UT_ISW(int m, int n, double *A, int lda, double *T, int ldt, double *R,
                            int ldr) {
    o_geqr3_R_form(1, 1, m, n, A, lda, T, ldt, R, ldr);
    return 0;
}

/* The dlarft3 family: ut / isw as above (qr2_dlarft3_UT.c:3-43,
 * qr2_dlarft3_ISW.c:3-39, qr2_dlarft3_UT_ISW.c:3-39); v2t = 1 builds each
 * left half from V^T V instead of recursing (qr2_dlarft3_ISW_V2T.c:3-43,
 * qr2_dlarft3_ISW_V2T_UT.c:3-43: V11^T into the strict upper part, dV2N,
 * the Gram of the rows below, then dN2T or the 1/tau diagonal). */
static void o_larft3_form(int ut, int isw, int v2t, int m, int n, double *A, int lda, double *T,
                          int ldt, double *tau) {
    if (n <= 1) {
        T[0] = ut ? 1.0 / tau[0] : tau[0];
        return;
    }
    const int n1 = n / 2, n2 = n - n1;
    double *A21 = A + n1, *A31 = A + n, *A22 = A + n1 + (size_t)n1 * lda, *A32 = A + n + (size_t)n1 * lda;
    double *T12 = T + (size_t)n1 * ldt, *T22 = T + n1 + (size_t)n1 * ldt;
    if (v2t) {
        for (int c = 0; c < n1; c++)
            for (int r = 0; r < c; r++) IDX(T, r, c, ldt) = IDX(A, c, r, lda);
        ora_qr2_aux_dV2N(n1, T, ldt);
        ora_dsyrk('U', 'T', n1, m - n1, 1.0, A + n1, lda, 1.0, T, ldt);
        if (ut)
            for (int c = 0; c < n1; c++) IDX(T, c, c, ldt) = 1.0 / tau[c];
        else
            ora_qr2_aux_dN2T(n1, tau, T, ldt);
    } else {
        o_larft3_form(ut, 0, 0, m, n1, A, lda, T, ldt, tau);
    }
    o_larft3_form(ut, isw, v2t, m - n1, n2, A22, lda, T22, ldt, tau + n1);
    if (isw) return;
    for (int r = 0; r < n1; r++)
        for (int c = 0; c < n2; c++) IDX(T12, r, c, ldt) = IDX(A21, c, r, lda);
    ora_dtrmm('R', 'L', 'N', 'U', n1, n2, 1.0, A22, lda, T12, ldt);
    ora_dgemm('T', 'N', n1, n2, m - n, 1.0, A31, lda, A32, lda, 1.0, T12, ldt);
    if (!ut) {
        ora_dtrmm('L', 'U', 'N', 'N', n1, n2, -1.0, T, ldt, T12, ldt);
        ora_dtrmm('R', 'U', 'N', 'N', n1, n2, 1.0, T22, ldt, T12, ldt);
    }
}

int ora_qr2_dlarft3_UT(int m, int n, double *A, int lda, double *T, int ldt, double *tau) {
    o_larft3_form(1, 0, 0, m, n, A, lda, T, ldt, tau);
    return 0;
}
int ora_qr2_dlarft3_ISW(int m, int n, double *A, int lda, double *T, int ldt, double *tau) {
    o_larft3_form(0, 1, 0, m, n, A, lda, T, ldt, tau);
    return 0;
}
int ora_qr2_dlarft3_UT_ISW(int m, int n, double *A, int lda, double *T, int ldt, double *tau) {
    o_larft3_form(1, 1, 0, m, n, A, lda, T, ldt, tau);
    return 0;
}
int ora_qr2_dlarft3_ISW_V2T(int m, int n, double *A, int lda, double *T, int ldt, double *tau) {
    o_larft3_form(0, 1, 1, m, n, A, lda, T, ldt, tau);
    return 0;
}
int ora_qr2_dlarft3_ISW_V2T_UT(int m, int n, double *A, int lda, double *T, int ldt, double *tau) {
    o_larft3_form(1, 1, 1, m, n, A, lda, T, ldt, tau);
    return 0;
}

/* The dorgqr3 family: ut solves with S11 where the T form multiplies by T11
 * (qr2_dorgqr3_UT.c:3-53); aux = 1 hands each left half to dVT2Q / dVS2Q
 * after copying its T (S) block into Q's upper part (qr2_dorgqr3_VT2Q.c:3-54,
 * qr2_dorgqr3_VS2Q.c:3-54).  form(0, 0) is ora_qr2_dorgqr3. */
static void o_orgqr3_form(int ut, int aux, int m, int n, double *Q, int ldq, double *T, int ldt,
                          double *tau) {
    if (n <= 1) {
        for (int r = 1; r < m; r++) Q[r] *= -tau[0];
        Q[0] = 1.0 - tau[0];
        return;
    }
    const int n1 = n / 2, n2 = n - n1;
    double *Q21 = Q + n1, *Q12 = Q + (size_t)n1 * ldq, *Q22 = Q + n1 + (size_t)n1 * ldq;
    o_orgqr3_form(ut, aux, m - n1, n2, Q22, ldq, T + n1 + (size_t)n1 * ldt, ldt, tau + n1);
    ora_dgemm('T', 'N', n1, n2, m - n1, 1.0, Q21, ldq, Q22, ldq, 0.0, Q12, ldq);
    if (ut) ora_dtrsm('L', 'U', 'N', 'N', n1, n2, 1.0, T, ldt, Q12, ldq);
    else ora_dtrmm('L', 'U', 'N', 'N', n1, n2, 1.0, T, ldt, Q12, ldq);
    ora_dgemm('N', 'N', m - n1, n2, n1, -1.0, Q21, ldq, Q12, ldq, 1.0, Q22, ldq);
    ora_dtrmm('L', 'L', 'N', 'U', n1, n2, -1.0, Q, ldq, Q12, ldq);
    if (aux) {
        if (T != Q) ora_dlacpy('U', n1, n1, T, ldt, Q, ldq);
        if (ut) ora_qr2_aux_dVS2Q(m, n1, Q, ldq);
        else ora_qr2_aux_dVT2Q(m, n1, Q, ldq);
    } else {
        o_orgqr3_form(ut, 0, m, n1, Q, ldq, T, ldt, tau);
    }
}

int ora_qr2_dorgqr3_UT(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau) {
    o_orgqr3_form(1, 0, m, n, Q, ldq, T, ldt, tau);
    return 0;
}
int ora_qr2_dorgqr3_VT2Q(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau) {
    o_orgqr3_form(0, 1, m, n, Q, ldq, T, ldt, tau);
    return 0;
}
int ora_qr2_dorgqr3_VS2Q(int m, int n, double *Q, int ldq, double *T, int ldt, double *tau) {
    o_orgqr3_form(1, 1, m, n, Q, ldq, T, ldt, tau);
    return 0;
}

/* qr3/src/qr2_aux_dV2N.c:3-30: upper(T) <- U U^T, U = I + strict upper of
 * T, by halves: T11 <- U11 U11^T + T12 T12^T, T12 <- T12 U22^T */
int ora_qr2_aux_dV2N(int n, double *T, int ldt) {
    if (n <= 1) {
        T[0] = 1.0;
        return 0;
    }
    const int n1 = n / 2, n2 = n - n1;
    double *T12 = T + (size_t)n1 * ldt, *T22 = T + n1 + (size_t)n1 * ldt;
    ora_qr2_aux_dV2N(n1, T, ldt);
    ora_dsyrk('U', 'N', n1, n2, 1.0, T12, ldt, 1.0, T, ldt);
    ora_dtrmm('R', 'U', 'T', 'U', n1, n2, 1.0, T22, ldt, T12, ldt);
    ora_qr2_aux_dV2N(n2, T22, ldt);
    return 0;
}

/* qr3/src/qr2_aux_dN2T.c:3-40: T from tau and N (strict upper part of
 * V^T V) by halves, T12 <- -T11 N12 T22 */
int ora_qr2_aux_dN2T(int n, double *tau, double *T, int ldt) {
    if (n <= 1) {
        T[0] = tau[0];
        return 0;
    }
    const int n1 = n / 2, n2 = n - n1;
    double *T12 = T + (size_t)n1 * ldt, *T22 = T + n1 + (size_t)n1 * ldt;
    ora_qr2_aux_dN2T(n1, tau, T, ldt);
    ora_qr2_aux_dN2T(n2, tau + n1, T22, ldt);
    ora_dtrmm('L', 'U', 'N', 'N', n1, n2, -1.0, T, ldt, T12, ldt);
    ora_dtrmm('R', 'U', 'N', 'N', n1, n2, 1.0, T22, ldt, T12, ldt);
    return 0;
}

/* tail shared by qr3/src/qr2_aux_dVT2Q.c:3-17 and qr2_aux_dVS2Q.c:3-17 once
 * the top block's upper part holds X = T V1^T: Q2 <- -V2 X, Q1 <- I - V1 X */
static void o_vx2q(int m, int n, double *Q, int ldq) {
    ora_dtrmm('R', 'U', 'N', 'N', m - n, n, -1.0, Q, ldq, Q + n, ldq);
    ora_qr2_aux_dmLUinA(n, Q, ldq);
    for (int k = 0; k < n; k++) IDX(Q, k, k, ldq) += 1.0;
}

int ora_qr2_aux_dVT2Q(int m, int n, double *Q, int ldq) {
    ora_qr2_aux_dULTinU(n, Q, ldq, Q, ldq);
    o_vx2q(m, n, Q, ldq);
    return 0;
}

int ora_qr2_aux_dVS2Q(int m, int n, double *Q, int ldq) {
    ora_qr2_aux_dUinvLTinU(n, Q, ldq, Q, ldq);
    o_vx2q(m, n, Q, ldq);
    return 0;
}

/* qr3/src/qr2_aux_dmLUinA.c:3-40: A <- -L U (L unit lower, U upper, both in
 * A), bottom-right first: A22 <- -L22 U22 - L21 U12, A12 <- -L11 U12,
 * A21 <- -L21 U11, then A11 */
int ora_qr2_aux_dmLUinA(int n, double *A, int lda) {
    if (n <= 1) {
        A[0] = -A[0];
        return 0;
    }
    const int n1 = n / 2, n2 = n - n1;
    double *A12 = A + (size_t)n1 * lda, *A21 = A + n1, *A22 = A + n1 + (size_t)n1 * lda;
    ora_qr2_aux_dmLUinA(n2, A22, lda);
    ora_dgemm('N', 'N', n2, n2, n1, -1.0, A21, lda, A12, lda, 1.0, A22, lda);
    ora_dtrmm('L', 'L', 'N', 'U', n1, n2, -1.0, A, lda, A12, lda);
    ora_dtrmm('R', 'U', 'N', 'N', n2, n1, -1.0, A, lda, A21, lda);
    ora_qr2_aux_dmLUinA(n1, A, lda);
    return 0;
}

/* qr3/src/qr2_aux_dV2tau.c:3-15 */
int ora_qr2_aux_dV2tau(int m, int n, double *A, int lda, double *tau) {
    for (int k = 0; k < n; k++) {
        double s = 1.0;
        for (int r = k + 1; r < m; r++) s += IDX(A, r, k, lda) * IDX(A, r, k, lda);
        tau[k] = 2.0 / s;
    }
    if (m == n && n > 0) tau[m - 1] = 0.0;
    return 0;
}

/* qr3/src/qr2_aux_dULTinU.c:3-38: upper(U) <- U L^T by halves:
 * U12 <- U12 L22^T + U11 L21^T, then U11 and U22 */
int ora_qr2_aux_dULTinU(int n, double *L, int ldl, double *U, int ldu) {
    if (n <= 1) return 0;
    const int n1 = n / 2, n2 = n - n1;
    double *U12 = U + (size_t)n1 * ldu, *U22 = U + n1 + (size_t)n1 * ldu;
    double *L21 = L + n1, *L22 = L + n1 + (size_t)n1 * ldl;
    ora_dtrmm('R', 'L', 'T', 'U', n1, n2, 1.0, L22, ldl, U12, ldu);
    ora_qr2_aux_dApUBTinA(n1, n2, U12, ldu, U, ldu, L21, ldl);
    ora_qr2_aux_dULTinU(n1, L, ldl, U, ldu);
    ora_qr2_aux_dULTinU(n2, L22, ldl, U22, ldu);
    return 0;
}

/* qr3/src/qr2_aux_dApUBTinA.c:46-96: A (m x n) += U B^T with U upper (m x m)
 * and B n x m, by halves of both dimensions */
int ora_qr2_aux_dApUBTinA(int m, int n, double *A, int lda, double *U, int ldu, double *B, int ldb) {
    if (m <= 0 || n <= 0) return 0;
    if (m <= 1 && n <= 1) {
        A[0] += U[0] * B[0];
        return 0;
    }
    const int m1 = m / 2, m2 = m - m1, n1 = n / 2, n2 = n - n1;
    double *A12 = A + (size_t)n1 * lda, *A21 = A + m1, *A22 = A + m1 + (size_t)n1 * lda;
    double *B12 = B + (size_t)m1 * ldb, *B21 = B + n1, *B22 = B + n1 + (size_t)m1 * ldb;
    double *U12 = U + (size_t)m1 * ldu, *U22 = U + m1 + (size_t)m1 * ldu;
    if (m1 > 0 && n1 > 0) ora_qr2_aux_dApUBTinA(m1, n1, A, lda, U, ldu, B, ldb);
    if (m1 > 0 && n1 > 0 && m2 > 0)
        ora_dgemm('N', 'T', m1, n1, m2, 1.0, U12, ldu, B12, ldb, 1.0, A, lda);
    if (m1 > 0 && n2 > 0) ora_qr2_aux_dApUBTinA(m1, n2, A12, lda, U, ldu, B21, ldb);
    if (m1 > 0 && n2 > 0 && m2 > 0)
        ora_dgemm('N', 'T', m1, n2, m2, 1.0, U12, ldu, B22, ldb, 1.0, A12, lda);
    if (m2 > 0 && n1 > 0) ora_qr2_aux_dApUBTinA(m2, n1, A21, lda, U22, ldu, B12, ldb);
    if (m2 > 0 && n2 > 0) ora_qr2_aux_dApUBTinA(m2, n2, A22, lda, U22, ldu, B22, ldb);
    return 0;
}

/* qr3/src/qr2_aux_dUinvLTinU.c:3-37: upper(U) <- U^-1 L^T, bottom-right
 * first: U12 <- U11^-1 (L21^T - U12 X22) with X22 the new U22 */
int ora_qr2_aux_dUinvLTinU(int n, double *L, int ldl, double *U, int ldu) {
    if (n <= 1) {
        U[0] = 1.0 / U[0];
        return 0;
    }
    const int n1 = n / 2, n2 = n - n1;
    double *U12 = U + (size_t)n1 * ldu, *U22 = U + n1 + (size_t)n1 * ldu;
    double *L21 = L + n1, *L22 = L + n1 + (size_t)n1 * ldl;
    ora_qr2_aux_dUinvLTinU(n2, L22, ldl, U22, ldu);
    ora_dtrmm('R', 'U', 'N', 'N', n1, n2, -1.0, U22, ldu, U12, ldu);
    for (int r = 0; r < n1; r++)
        for (int c = 0; c < n2; c++) IDX(U12, r, c, ldu) += IDX(L21, c, r, ldl);
    ora_dtrsm('L', 'U', 'N', 'N', n1, n2, 1.0, U, ldu, U12, ldu);
    ora_qr2_aux_dUinvLTinU(n1, L, ldl, U, ldu);
    return 0;
}

/* LAPACK dlarf as called by qr3/src/qr2_aux_dlarf_wrapper.c:3-8, v with
 * stride incv >= 1: C <- (I - tau v v^T) C ('L') or C (I - tau v v^T) ('R') */
int ora_qr2_aux_dlarf_wrapper(char side, int m, int n, double *V, int incv, double tau, double *C,
                              int ldc, double *work) {
    (void)work;
    if ((side != 'L' && side != 'R') || incv < 1) return -1;
    if (tau == 0.0 || m <= 0 || n <= 0) return 0;
    if (side == 'L') {
        for (int c = 0; c < n; c++) {
            double s = 0.0;
            for (int r = 0; r < m; r++) s += V[(size_t)r * incv] * IDX(C, r, c, ldc);
            s *= tau;
            for (int r = 0; r < m; r++) IDX(C, r, c, ldc) -= s * V[(size_t)r * incv];
        }
    } else {
        for (int r = 0; r < m; r++) {
            double s = 0.0;
            for (int c = 0; c < n; c++) s += IDX(C, r, c, ldc) * V[(size_t)c * incv];
            s *= tau;
            for (int c = 0; c < n; c++) IDX(C, r, c, ldc) -= s * V[(size_t)c * incv];
        }
    }
    return 0;
}

/* ---- checkers ------------------------------------------------------------ */

/* ||triu(Q^T Q - I)||_F, the half norm of
 * lila-c-lila_param/lila/src/lila_test_qq_orth_1.c:13-15 */
double ora_test_qq_orth_1(int m, int n, int i, const double *Q, int ldq) {
    double *W = (double *)xmalloc((size_t)n * n * sizeof(double));
    ora_dlaset('A', n, n, 0.0, 1.0, W, n);
    ora_dsyrk('U', 'T', n, m - i, 1.0, &IDX(Q, i, i, ldq), ldq, -1.0, W, n);
    double r = ora_dlange_f(n, n, W, n);
    free(W);
    return r;
}

/* full ||Q^T Q - I||_F */
double ora_test_qq_orth_full(int m, int n, int i, const double *Q, int ldq) {
    double *W = (double *)xmalloc((size_t)n * n * sizeof(double));
    ora_dlaset('A', n, n, 0.0, 1.0, W, n);
    ora_dgemm('T', 'N', n, n, m - i, 1.0, &IDX(Q, i, i, ldq), ldq, &IDX(Q, i, i, ldq), ldq,
              -1.0, W, n);
    double r = ora_dlange_f(n, n, W, n);
    free(W);
    return r;
}

/* ||Q triu(R) - A||_F / ||A||_F;
 * lila-c-lila_param/lila/src/lila_test_qr_repres_1.c:12-25 */
double ora_test_qr_repres_1(int m, int n, int i, const double *A, int lda,
                            const double *Q, int ldq, const double *R, int ldr) {
    int ml = m - i;
    const double *Aii = &IDX(A, i, i, lda);
    double nA = ora_dlange_f(ml, n, Aii, lda);
    double *W = (double *)xmalloc((size_t)ml * n * sizeof(double));
    ora_dlacpy('A', ml, n, &IDX(Q, i, i, ldq), ldq, W, ml);
    ora_dtrmm('R', 'U', 'N', 'N', ml, n, 1.0, &IDX(R, i, i, ldr), ldr, W, ml);
    for (int c = 0; c < n; c++)
        for (int r = 0; r < ml; r++) IDX(W, r, c, ml) -= IDX(Aii, r, c, lda);
    double r = ora_dlange_f(ml, n, W, ml) / nA;
    free(W);
    return r;
}

/* Apply H^T (k reflectors from V,T stored in R/T) to a copy of As and
 * compare: lower part against 0, upper part against R;
 * lila-c-lila_param/lila/src/lila_test_r_repres_2.c:14-41 */
double ora_test_r_repres_2(int m, int n, int k, int i, int mt,
                           const double *As, int lda, const double *R, int ldr,
                           const double *T, int ldt) {
    int ml = m - i;
    double *RR = (double *)xmalloc((size_t)m * n * sizeof(double));
    ora_dlacpy('A', ml, n, &IDX(As, i, i, lda), lda, RR + i, m);
    double nA = ora_dlange_f(ml, n, &IDX(As, i, i, lda), lda);
    o_apply_qt(m, n, k, i, mt, R, ldr, T, ldt, RR + i, m);
    double s1 = 0.0, s2 = 0.0;
    for (int c = 0; c < n; c++)
        for (int r = 0; r < ml; r++) {
            double v = RR[(size_t)i + r + (size_t)c * m];
            if (r > c) s1 += v * v;
            else if (r < n) { double d = IDX(R, i + r, i + c, ldr) - v; s2 += d * d; }
        }
    free(RR);
    return o_dlapy2(sqrt(s1), sqrt(s2)) / nA;
}

/* ---- generators ---------------------------------------------------------- */
static uint64_t sm64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void ora_fill_uniform(unsigned long long seed, int m, int n, double *A, int ld,
                      double lo, double hi) {
    uint64_t s = seed;
    for (int c = 0; c < n; c++)
        for (int r = 0; r < m; r++) {
            double u = (double)(sm64(&s) >> 11) * 0x1.0p-53;
            IDX(A, r, c, ld) = lo + (hi - lo) * u;
        }
}

/* element e (column-major order) uses draws 2*(e/2), 2*(e/2)+1: even e takes
 * r*cos(2 pi u2), odd e takes r*sin(2 pi u2), r = sqrt(-2 ln u1),
 * u1 = (x+1) 2^-53 in (0,1]. */
void ora_fill_normal(unsigned long long seed, int m, int n, double *A, int ld) {
    uint64_t s = seed;
    long long total = (long long)m * n;
    double z1 = 0.0;
    for (long long e = 0; e < total; e++) {
        int r = (int)(e % m), c = (int)(e / m);
        if ((e & 1) == 0) {
            double u1 = (double)((sm64(&s) >> 11) + 1) * 0x1.0p-53;
            double u2 = (double)(sm64(&s) >> 11) * 0x1.0p-53;
            double rad = sqrt(-2.0 * log(u1));
            double th = 6.283185307179586 * u2;
            IDX(A, r, c, ld) = rad * cos(th);
            z1 = rad * sin(th);
        } else {
            IDX(A, r, c, ld) = z1;
        }
    }
}
