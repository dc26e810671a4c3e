/*
 * flops.c -- exact operation counts of the routines of the recursive QR
 * library, with the formulas of the reference's flop-count library
 * (qr3/flops/, declared in qr3/flops/flops.h): a closed form per routine and,
 * for the recursive ones, a `_check` form that adds up the operation counts
 * of the recursion step by step, the pair the reference uses to validate
 * its closed forms.  Host arithmetic in 64-bit integers (the reference
 * multiplies some counts in `int`; the results agree wherever those products
 * fit in 32 bits).  flops_geqr3's closed form equals flops_geqr3_check when n
 * is a power of two (the sizes it was derived for); the others agree for all
 * sizes.
 */
#include "../../include/lila.h"

typedef long long i64;

/* qr3/flops/flops_gemm.c:3-15 */
long int flops_gemm(int m, int n, int k) { return (long int)(2 * (i64)m * n * k); }

/* qr3/flops/flops_trmm.c:3-20: triangular factor of order m ('L') or n ('R') */
long int flops_trmm(char S, int m, int n) {
    if (S == 'L') return (long int)((i64)n * m * m);
    if (S == 'R') return (long int)((i64)m * n * n);
    return 0;
}

/* qr3/flops/flops_syrk.c:3-14 */
long int flops_syrk(int n, int k) { return (long int)((i64)k * n * n); }

/* qr3/flops/flops_larft.c:5-22 */
long int flops_larft(int m, int n) {
    const i64 M = m, N = n;
    return (long int)((6 * M * N * N - 2 * N * N * N - 6 * M * N + 3 * N * N - N) / 6);
}

/* qr3/flops/flops_lapack_larfg.c:3-13 */
long int flops_lapack_larfg(int m) { return (long int)(3 * (i64)m + 5); }

/* qr3/flops/flops_lapack_larf.c:3-14 */
long int flops_lapack_larf(int m, int n) { return (long int)(4 * (i64)m * n - n); }

/* qr3/flops/flops_lapack_geqr2.c:3-17 */
long int flops_lapack_geqr2(int m, int n) {
    const i64 M = m, N = n;
    return (long int)((6 * M * N * N - 2 * N * N * N + 3 * M * N + 17 * N) / 3);
}

/* qr3/flops/flops_lapack_org2r.c:3-77: Q1 (the first k columns) and Q2 */
long int flops_lapack_org2r(int m, int n, int k) {
    const i64 M = m, N = n, K = k;
    i64 f = (6 * (M - K) * K + 4 * K * K - 3 * (M - K) - 1) * K / 3;
    f += (4 * M - 2 * K + 1) * K * (N - K);
    return (long int)f;
}

/* qr3/flops/flops_lapack_org2r_check.c:3-51: one larf and one scaling per
 * column, from the last reflector back */
long int flops_lapack_org2r_check(int m, int n, int k) {
    if (n == 1) return m;
    i64 ml = (i64)m - k + 1, nl = (i64)n - k + 1;
    i64 f = 4 * ml * (nl - 1) - (nl - 1) + ml;
    while (nl < n) {
        ml++;
        nl++;
        f += 4 * ml * (nl - 1) - (nl - 1);
        f += ml;
    }
    return (long int)f;
}

/* qr3/flops/flops_geqr3.c:3-67 */
long int flops_geqr3(int m, int n) {
    const i64 M = m, N = n;
    return (long int)((18 * M * N * N - 5 * N * N * N + 35 * N) / 6);
}

/* qr3/flops/flops_geqr3_check.c:3-44: the steps of qr3/src/qr2_dgeqr3_R.c
 * (larfg at a single column; the update of the right half: trmm, gemm,
 * trmm (with the extra of a sequence of larf), trmm, gemm, trmm; the right
 * half; T12: trmm, gemm, trmm, trmm) */
long int flops_geqr3_check(int m, int n) {
    if (n == 1) return flops_lapack_larfg(m);
    const i64 n1 = n / 2, n2 = n - n1, M = m, N = n;
    i64 f = flops_geqr3_check(m, (int)n1);
    f += n1 * n1 * n2;
    f += 2 * n1 * n2 * (M - n1);
    f += (n1 - 1) * n1 * n2;
    f += n1 * n2;
    f += 2 * n1 * n2 * (M - n1);
    f += n1 * n1 * n2;
    f += flops_geqr3_check((int)(M - n1), (int)n2);
    f += n1 * n2 * n2;
    f += 2 * n1 * n2 * (M - N);
    f += n1 * n1 * n2;
    f += n1 * n2 * n2;
    return (long int)f;
}

/* qr3/flops/flops_dlarft3.c:3-37 */
long int flops_dlarft3(int m, int n) {
    if (n <= 1) return 0;
    const i64 n1 = n / 2, n2 = n - n1, M = m, N = n;
    i64 f = flops_dlarft3(m, (int)n1) + flops_dlarft3((int)(M - n1), (int)n2);
    f += n1 * n2 * n2 + 2 * n1 * n2 * M - 2 * n1 * n2 * N + n1 * n1 * n2 + n1 * n2 * n2;
    return (long int)f;
}

/* qr3/flops/flops_dlarft3_check.c:3-39: T12 = trmm, gemm over the rows below
 * both halves, trmm, trmm */
long int flops_dlarft3_check(int m, int n) {
    if (n <= 1) return 0;
    const i64 n1 = n / 2, n2 = n - n1, M = m, N = n;
    i64 f = flops_dlarft3_check(m, (int)n1) + flops_dlarft3_check((int)(M - n1), (int)n2);
    f += n1 * n2 * n2 + 2 * n1 * n2 * (M - N) + n1 * n1 * n2 + n1 * n2 * n2;
    return (long int)f;
}

/* qr3/flops/flops_qr2_dorgqr.c:3-31 */
long int flops_qr2_dorgqr(int m, int n) {
    const i64 M = m, N = n;
    i64 f = (6 * (M - N) * N + 4 * N * N - 3 * (M - N) - 1) * N / 3;
    f -= N * (N - 1) / 2;
    return (long int)f;
}

/* qr3/flops/flops_qr2_dorgqr_check.c:3-39: the steps of
 * qr3/src/qr2_dorgqr3.c (a column: m; the right half, gemm, trmm with the
 * extra of a larf sequence, the diagonal, gemm, trmm, the left half) */
long int flops_qr2_dorgqr_check(int m, int n) {
    if (n <= 1) return m;
    const i64 n1 = n / 2, n2 = n - n1, M = m;
    i64 f = flops_qr2_dorgqr_check((int)(M - n1), (int)n2);
    f += 2 * n1 * n2 * (M - n1);
    f += (n1 - 1) * n1 * n2;
    f += n1 * n2;
    f += 2 * n1 * n2 * (M - n1);
    f += n1 * n1 * n2;
    f += flops_qr2_dorgqr_check(m, (int)n1);
    return (long int)f;
}

/* qr3/flops/flops_mLUinA.c:3-13 */
long int flops_mLUinA(int n) {
    const i64 N = n;
    return (long int)(2 * (N * N * N - N) / 3);
}

/* qr3/flops/flops_mLUinA_check.c:3-33: one level of qr3/src/qr2_aux_dmLUinA.c
 * with the closed form for the halves */
long int flops_mLUinA_check(int n) {
    if (n <= 1) return 0;
    const int n1 = n / 2, n2 = n - n1;
    return flops_mLUinA(n2) + flops_gemm(n2, n2, n1) + flops_trmm('L', n1, n2) +
           flops_trmm('R', n2, n1) + flops_mLUinA(n1);
}

/* qr3/flops/flops_V2N.c:3-13 */
long int flops_V2N(int n) {
    const i64 N = n;
    return (long int)((N * N * N - N) / 3);
}

/* qr3/flops/flops_V2N_check.c:3-33 */
long int flops_V2N_check(int n) {
    if (n <= 1) return 0;
    const int n1 = n / 2, n2 = n - n1;
    return flops_V2N(n1) + flops_syrk(n1, n2) + flops_trmm('R', n1, n2) + flops_V2N(n2);
}

/* qr3/flops/flops_ApUBTinA.c:3-15 */
long int flops_ApUBTinA(int m, int n) { return (long int)((i64)m * m * n + (i64)m * n); }

/* qr3/flops/flops_ApUBTinA_check.c:3-40: the recursion of
 * qr3/src/qr2_aux_dApUBTinA.c (2 at a single entry) */
long int flops_ApUBTinA_check(int m, int n) {
    if (m <= 1 && n <= 1) return 2;
    const int m1 = m / 2, m2 = m - m1, n1 = n / 2, n2 = n - n1;
    i64 f = 0;
    if (m1 > 0 && n1 > 0) f += flops_ApUBTinA_check(m1, n1);
    if (m1 > 0 && n1 > 0 && m2 > 0) f += flops_gemm(m1, n1, m2);
    if (m1 > 0 && n2 > 0) f += flops_ApUBTinA_check(m1, n2);
    if (m1 > 0 && n2 > 0 && m2 > 0) f += flops_gemm(m1, n2, m2);
    if (m2 > 0 && n1 > 0) f += flops_ApUBTinA_check(m2, n1);
    if (m2 > 0 && n2 > 0) f += flops_ApUBTinA_check(m2, n2);
    return (long int)f;
}

/* qr3/flops/flops_ULTinU.c:3-13 */
long int flops_ULTinU(int n) {
    const i64 N = n;
    return (long int)((2 * N * N * N + 3 * N * N - 5 * N) / 6);
}

/* qr3/flops/flops_ULTinU_check.c:3-31 */
long int flops_ULTinU_check(int n) {
    if (n <= 1) return 0;
    const int n1 = n / 2, n2 = n - n1;
    return flops_trmm('R', n1, n2) + flops_ApUBTinA(n1, n2) + flops_ULTinU_check(n1) +
           flops_ULTinU_check(n2);
}

/* qr3/flops/flops_N2T.c:3-13 */
long int flops_N2T(int n) {
    const i64 N = n;
    return (long int)((N * N * N - N) / 3);
}

/* qr3/flops/flops_N2T_check.c:3-31 */
long int flops_N2T_check(int n) {
    if (n <= 1) return 0;
    const int n1 = n / 2, n2 = n - n1;
    return flops_N2T_check(n1) + flops_N2T_check(n2) + flops_trmm('L', n1, n2) +
           flops_trmm('R', n1, n2);
}

/* qr3/flops/flops_VT2Q.c:3-14 */
long int flops_VT2Q(int m, int n) {
    const i64 M = m, N = n;
    return (long int)(M * N * N + (N * N - N) / 2);
}

/* qr3/flops/flops_VT2Q_check.c:3-19: the steps of qr3/src/qr2_aux_dVT2Q.c */
long int flops_VT2Q_check(int m, int n) {
    return flops_ULTinU(n) + flops_trmm('R', m - n, n) + flops_mLUinA(n) + (long int)n;
}
