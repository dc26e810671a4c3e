// small32.h -- one-wave factorizations of 32 x 32 matrices held in LDS
// (leading dimension 34), shared by the Householder panel and the batched
// least-squares kernels.  Each routine runs on one wave; lane c (c < 32)
// keeps column c in registers and reads other columns' entries with
// v_readlane, so all 64 lanes of the wave must be active.
#ifndef LILA_SMALL32_H
#define LILA_SMALL32_H
#include <hip/hip_runtime.h>

namespace small32 {

constexpr int PNB = 32;  // matrix order
constexpr int LDM = 34;  // leading dimension in LDS (even: 16-byte pairs)

// the lane index as a value the compiler cannot hoist: the lane masks and
// addresses derived from it are then formed where they are used instead of
// once at kernel entry (where, held for the whole kernel, they would crowd
// out the registers of the routine's own loops)
__device__ __forceinline__ int opaque_lane() {
  int t = threadIdx.x & 63;
  asm volatile("" : "+v"(t));
  return t;
}

// LDS ordering between the lanes of one wave (wave-only 32 x 32 routines)
__device__ __forceinline__ void wfence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// v from lane l of the wave (l a compile-time constant after unrolling:
// two v_readlane_b32, the result lives in SGPRs)
__device__ __forceinline__ double rdlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

typedef double d4s_t __attribute__((ext_vector_type(4)));

// one 16 x 16 x 16 product on v_mfma_f64_16x16x4f64 by one wave: returns the
// lane's four entries of A B, A(i, k) at A[i * ars + k * acs], B(k, j) at
// B[k * brs + j * bcs]; entry v of lane l is (l / 16 + 4 v, l % 16)
__device__ __forceinline__ d4s_t mm16(const double *A, int ars, int acs, const double *B, int brs,
                                      int bcs) {
  const int lane = opaque_lane();
  d4s_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < 4; ks++) {
    const int k = 4 * ks + (lane >> 4);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A[(lane & 15) * ars + k * acs],
                                               B[k * brs + (lane & 15) * bcs], acc, 0, 0, 0);
  }
  return acc;
}

// upper Cholesky of the 16 x 16 block of M at (o, o) in place (lower part of
// the block zeroed), 1 / diag(R) into D[o..o+15]; lane c (c < 16) keeps
// column c in registers, row broadcasts by v_readlane (all lanes run).
// ok, dmin, dmax accumulate the pivot test and the spread of R's diagonal
// over the first nb columns of the 32 x 32 matrix.
__device__ __forceinline__ void chol16(double *M, double *D, int o, int nb, bool &ok, double &dmin,
                                       double &dmax) {
  const int lane = opaque_lane(), c = lane & 15;
  nb = __builtin_amdgcn_readfirstlane(nb);  // scalar branches (see lu16)
  double a[16];
#pragma unroll
  for (int r = 0; r < 16; r++) a[r] = M[(o + r) * LDM + o + c];
  double myinv = 1.0;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const double d = rdlane(a[j], j);
    // a pivot must be positive and finite: an overflowed Gram (inf) or a
    // NaN from one fails here, not only through the spread test
    ok = ok && (d > 0.0) && (d <= 1.7976931348623157e308);
    // 1 / sqrt(d) from the hardware estimate and two Newton steps
    // y <- y (3 - d y^2) / 2 (quadratic convergence: full precision), then
    // sqrt(d) = d / sqrt(d); both within an ulp or two of the rounded values
    const double dd = (d > 0.0) ? d : 1.0;
    double pinv = __builtin_amdgcn_rsq(dd);
    pinv = pinv * (1.5 - 0.5 * dd * pinv * pinv);
    pinv = pinv * (1.5 - 0.5 * dd * pinv * pinv);
    const double p = dd * pinv;
    if (o + j < nb) {
      dmax = fmax(dmax, p);
      dmin = fmin(dmin, p);
    }
    myinv = (c == j) ? pinv : myinv;
    const double rj = (c == j) ? p : a[j] * pinv;
    a[j] = rj;
#pragma unroll
    for (int r = j + 1; r < 16; r++) a[r] -= rdlane(rj, r) * rj;
  }
  if (lane < 16) {
#pragma unroll
    for (int r = 0; r < 16; r++) M[(o + r) * LDM + o + c] = (r <= c) ? a[r] : 0.0;
    D[o + c] = myinv;
  }
  wfence();
}

// X (block origin, ld LDM) = inverse of the upper 16 x 16 block of R at
// (o, o), D = 1 / diag; lanes 0..15 form the columns (back substitution,
// column in registers)
__device__ __forceinline__ void inv16(const double *R, const double *D, int o, double *X) {
  const int lane = opaque_lane(), c = lane & 15;
  double x[16];
#pragma unroll
  for (int i = 15; i >= 0; i--) {
    double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
    for (int q = i + 1; q < 16; q++) {
      const double t = R[(o + i) * LDM + o + q] * x[q];
      if (q & 1) acc1 += t; else acc0 += t;
    }
    const double di = D[o + i];
    x[i] = (i < c) ? -di * (acc0 + acc1) : (i == c ? di : 0.0);
  }
  if (lane < 16) {
#pragma unroll
    for (int i = 0; i < 16; i++) X[i * LDM + c] = x[i];
  }
}

// one wave: upper Cholesky of M (32 x 32, ld LDM, upper part read; rows
// and columns >= nb hold the identity) in place, lower part zeroed, and the
// reciprocals of R's diagonal into D, in 16 x 16 blocks: R11 = chol(A11),
// R12 = R11^-T A12 (R11^-1 formed in the lower-left block, which is free,
// then one MFMA product), A22 -= R12^T R12 (MFMA), R22 = chol(A22).  The
// pivots are those of the column-by-column algorithm.  Returns false if a
// pivot is not positive or R's first nb diagonal entries spread by more
// than 1e5.
__device__ __forceinline__ bool wave_chol32(double *M, double *D, int nb, bool spread = true) {
  const int lane = opaque_lane();
  bool ok = true;
  double dmin = 1e308, dmax = 0.0;
  chol16(M, D, 0, nb, ok, dmin, dmax);
  inv16(M, D, 0, M + 16 * LDM);  // R11^-1 in the lower-left block
  wfence();
  const d4s_t r12 = mm16(M + 16 * LDM, 1, LDM, M + 16, LDM, 1);  // (R11^-1)^T A12
#pragma unroll
  for (int v = 0; v < 4; v++) M[((lane >> 4) + 4 * v) * LDM + 16 + (lane & 15)] = r12[v];
  wfence();
  const d4s_t s22 = mm16(M + 16, 1, LDM, M + 16, LDM, 1);  // R12^T R12
#pragma unroll
  for (int v = 0; v < 4; v++) M[(16 + (lane >> 4) + 4 * v) * LDM + 16 + (lane & 15)] -= s22[v];
  wfence();
  chol16(M, D, 16, nb, ok, dmin, dmax);
#pragma unroll
  for (int e = lane; e < 256; e += 64) M[(16 + e / 16) * LDM + e % 16] = 0.0;
  wfence();
  return ok && (!spread || dmin > 1e-5 * dmax);
}

// 1 / x for the pivots of the sign-chosen LU (|x| = |t| + 1 >= 1, finite):
// the hardware estimate refined by two Newton steps, within an ulp of the
// rounded quotient (the IEEE division sequence is twice as long a chain)
__device__ __forceinline__ double recip(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

// sign-chosen LU (see wave_lu_sign32) of the 16 x 16 block of M at (o, o)
// with the signs left unapplied: U above the diagonal as eliminated, the
// pivots on it, L below; S[o..], D[o..] = signs, 1 / pivots
__device__ __forceinline__ void lu16(double *M, double *S, double *D, int o, int nb, bool &ok) {
  const int lane = opaque_lane(), c = lane & 15;
  double a[16];
#pragma unroll
  for (int r = 0; r < 16; r++) a[r] = M[(o + r) * LDM + o + c];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    // Steps k >= nb (identity columns: t = 1, zeros below) run too, with
    // their effect masked by act, so that no step is a branch: a branch
    // makes the compiler move a[] between register sets at every step.
    const bool act = o + k < nb;
    const double t = rdlane(a[k], k);
    // the rule |1 - t| < |1 + t| (S = -1) holds exactly when t > 0; the
    // pivot S t - 1 is then -(1 + |t|) either way, so the chain from t to
    // the multiplier is one add, the reciprocal and a sign
    const bool flip = t > 0.0;
    if (act && fabs(t) > 1.0 - 1e-10) ok = false;
    const double sk = flip ? -1.0 : 1.0;
    const double ap = 1.0 + fabs(t);
    const double piv = -ap;
    const double rp = recip(ap);
    const double pinv = -rp;
    const double f = act ? (flip ? rp : -rp) : 0.0;  // S / pivot
    // the step's sign and reciprocal pivot are the same on every lane:
    // stored by all (column k reads them back at the end); past nb the
    // identity's 1, 1
    S[o + k] = act ? sk : 1.0;
    D[o + k] = act ? pinv : 1.0;
    // trailing columns: a(r, c) -= L(r, k) U(k, c) with L(r, k) = a(r, k) f,
    // one FMA per row with f U(k, c) formed once; column k itself stays
    // unscaled until the end (finished columns: fu = 0)
    const double fu = (c > k) ? f * a[k] : 0.0;
#pragma unroll
    for (int r = k + 1; r < 16; r++) a[r] = fma(-rdlane(a[r], k), fu, a[r]);
    a[k] = (act && c == k) ? piv : a[k];
  }
  wfence();
  // write-out: U above the diagonal with the deferred sign S[c] of its
  // column, the pivot on it, L = column * f_c below it; with o > 0 also the
  // rows 0..o-1 of these columns (U12 of the 32 x 32 matrix, formed before
  // these signs were known) take their signs
  if (lane < 16) {
    const double mys = S[o + c], myf = mys * D[o + c];
#pragma unroll
    for (int r = 0; r < 16; r++)
      M[(o + r) * LDM + o + c] = (r < c) ? a[r] * mys : (r == c ? a[r] : a[r] * myf);
    if (o > 0) {
      double u[16];
#pragma unroll
      for (int r = 0; r < 16; r++) u[r] = M[r * LDM + o + c];
#pragma unroll
      for (int r = 0; r < 16; r++) M[r * LDM + o + c] = u[r] * mys;
    }
  }
  wfence();
}

// one wave: sign-chosen LU of M (Q1 on entry, identity beyond nb) in place,
// the column rule of lila-c/lila_dorgh2_3.c:8-21: at step k, S_k = -1 when
// |1 - t| < |1 + t| for the diagonal t (column k negated), the pivot is
// S_k t - 1, the column below is divided by it and the trailing block gets
// the rank-1 update.  The negation of column k is carried as a sign (the
// steps before k are linear in column k) and applied to its U entries at the
// end.  In 16 x 16 blocks: LU of A11; L21 = A21 V^-1 with V = A11's U as
// eliminated and diagonal pivot * sign (lanes 0..15, one row each) while
// U12 = L11^-1 A12 (lanes 16..31, one column each); A22 -= L21 U12 (MFMA);
// LU of A22 -- the same pivots and signs as the column-by-column order.  U
// (upper, diagonal the pivots) and L (strict lower) overwrite M, S gets the
// signs and D the reciprocal pivots.  Returns false when some |t| > 1 -
// 1e-10 (tau within 1e-10 of 0 or 2).
__device__ __forceinline__ bool wave_lu_sign32(double *M, double *S, double *D, int nb) {
  const int lane = opaque_lane();
  bool ok = true;
  lu16(M, S, D, 0, nb, ok);
  if (lane < 32) {
    // one forward substitution per lane, both kinds in the same instruction
    // stream (no divergent halves): lanes 0..15 a row of L21 = A21 V^-1 (V =
    // A11's U with the signs S1 applied above the diagonal and the pivot
    // times its sign on it: z[t] = (S[t] a[t] - sum_k z[k] U(k, t)) D[t]),
    // lanes 16..31 a column of U12 = L11^-1 A12 (L11 unit lower: z[t] =
    // a[t] - sum_k L(t, k) z[k])
    const bool isl = lane < 16;
    const int i = lane & 15;
    const int sk = isl ? LDM : 1, st = isl ? 1 : LDM;  // coefficient (k, t) at k sk + t st
    double z[16];
#pragma unroll
    for (int t = 0; t < 16; t++) z[t] = isl ? M[(16 + i) * LDM + t] * S[t] : M[t * LDM + 16 + i];
#pragma unroll
    for (int t = 0; t < 16; t++) {
      double acc = z[t];
#pragma unroll
      for (int k = 0; k < t; k++) acc -= z[k] * M[k * sk + t * st];
      z[t] = isl ? acc * D[t] : acc;
    }
#pragma unroll
    for (int t = 0; t < 16; t++) {
      if (isl) M[(16 + i) * LDM + t] = z[t];
      else M[t * LDM + 16 + i] = z[t];
    }
  }
  wfence();
  const d4s_t s22 = mm16(M + 16 * LDM, LDM, 1, M + 16, LDM, 1);  // L21 U12
#pragma unroll
  for (int v = 0; v < 4; v++) M[(16 + (lane >> 4) + 4 * v) * LDM + 16 + (lane & 15)] -= s22[v];
  wfence();
  lu16(M, S, D, 16, nb, ok);  // also applies S2 to U12
  return ok;
}

// X = R^-1 for upper R (ld LDM, identity beyond the used order; whatever R
// holds below its diagonal is not read) with D = 1 / diag(R), by one whole
// wave, in 16 x 16 blocks: R = [A B; 0 C] gives R^-1 = [A^-1, -A^-1 B C^-1;
// 0, C^-1].  Lanes 0..15 form the columns of A^-1 and lanes 16..31 those of
// C^-1 at the same time (back substitution, the column in registers, fully
// unrolled: 16 sequential steps instead of 32); the off-diagonal block is two
// 16 x 16 x 16 products on v_mfma_f64_16x16x4f64.
__device__ __forceinline__ void inv_upper32(const double *R, const double *D, double *X) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  const int lane = opaque_lane(), c = lane & 15, b = (lane & 16);
  if (lane < 32) {
    double x[16];
#pragma unroll
    for (int i = 15; i >= 0; i--) {
      double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
      for (int q = i + 1; q < 16; q++) {
        const double t = R[(b + i) * LDM + b + q] * x[q];
        if (q & 1) acc1 += t; else acc0 += t;
      }
      const double di = D[b + i];
      x[i] = (i < c) ? -di * (acc0 + acc1) : (i == c ? di : 0.0);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
      X[(b + i) * LDM + b + c] = x[i];
      if (b == 0) X[(16 + i) * LDM + c] = 0.0;  // lower-left block
    }
  }
  wfence();
  // T = B C^-1 into the upper-right block, then X12 = -A^-1 T in place
  d4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < 4; ks++) {
    const int k = 4 * ks + (lane >> 4);
    t = __builtin_amdgcn_mfma_f64_16x16x4f64(R[c * LDM + 16 + k], X[(16 + k) * LDM + 16 + c], t, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; v++) X[((lane >> 4) + 4 * v) * LDM + 16 + c] = t[v];
  wfence();
  d4 u = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < 4; ks++) {
    const int k = 4 * ks + (lane >> 4);
    u = __builtin_amdgcn_mfma_f64_16x16x4f64(X[c * LDM + k], X[k * LDM + 16 + c], u, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; v++) X[((lane >> 4) + 4 * v) * LDM + 16 + c] = -u[v];
  wfence();
}

// C = alpha A B (+ C when acc) for 32 x 32 blocks, all four waves of a 256-thread workgroup (wave w
// the 16 x 16 output tile (w & 1, w >> 1)): A(i, k) at A[i ars + k acs] (0
// for i > k when amask: an upper factor whose lower part holds something
// else), B(k, j) at B[k brs + j bcs], C row-major with ld LDM.  C may alias
// A or B: every read precedes the first write.
__device__ __forceinline__ void wg_mm32(const double *A, int ars, int acs, bool amask, const double *B, int brs, int bcs,
                        double *C, double alpha, bool acc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ti = wave & 1, tj = wave >> 1, i = ti * 16 + (lane & 15), j = tj * 16 + (lane & 15);
  d4s_t r = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < 8; ks++) {
    const int k = 4 * ks + (lane >> 4);
    const double av = (amask && i > k) ? 0.0 : A[i * ars + k * acs];
    r = __builtin_amdgcn_mfma_f64_16x16x4f64(av, B[k * brs + j * bcs], r, 0, 0, 0);
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < 4; v++) {
    double *c = C + (ti * 16 + (lane >> 4) + 4 * v) * LDM + tj * 16 + (lane & 15);
    *c = acc ? *c + alpha * r[v] : alpha * r[v];
  }
  __syncthreads();
}

}  // namespace small32

#endif
