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

// one wave: upper Cholesky of M (32 x 32, ld LDM, upper part read; rows
// and columns >= nb hold the identity) in place, lower part zeroed, and the
// reciprocals of R's diagonal into D.  Lane c keeps column c in registers;
// step j scales row j by 1/R(j,j) (LAPACK dpotf2's scaling) and updates
// A(r,c) -= R(j,r) R(j,c) with R(j,r) read from lane r.  Returns false if a
// pivot is not positive or R's first nb diagonal entries spread by more
// than 1e5.
__device__ __forceinline__ bool wave_chol32(double *M, double *D, int nb) {
  const int lane = opaque_lane(), c = lane & 31;
  double a[PNB];
#pragma unroll
  for (int r = 0; r < PNB; r++) a[r] = M[r * LDM + c];
  bool ok = true;
  double dmax = 0.0, dmin = 1e308, myinv = 1.0;
#pragma unroll
  for (int j = 0; j < PNB; j++) {
    const double d = rdlane(a[j], j);
    ok = ok && (d > 0.0);
    // 1 / sqrt(d) from the hardware estimate and two Newton steps
    // y <- y (3 - d y^2) / 2 (quadratic convergence: full precision), then
    // sqrt(d) = d / sqrt(d); both within an ulp or two of the rounded values
    const double dd = (d > 0.0) ? d : 1.0;
    double pinv = __builtin_amdgcn_rsq(dd);
    pinv = pinv * (1.5 - 0.5 * dd * pinv * pinv);
    pinv = pinv * (1.5 - 0.5 * dd * pinv * pinv);
    const double p = dd * pinv;
    if (j < nb) {
      dmax = fmax(dmax, p);
      dmin = fmin(dmin, p);
    }
    if (c == j) myinv = pinv;
    const double rj = (c == j) ? p : a[j] * pinv;
    a[j] = rj;
#pragma unroll
    for (int r = j + 1; r < PNB; r++) a[r] -= rdlane(rj, r) * rj;
  }
  if (lane < PNB) {
#pragma unroll
    for (int r = 0; r < PNB; r++) M[r * LDM + c] = (r <= c) ? a[r] : 0.0;
    D[c] = myinv;
  }
  wfence();
  return ok && dmin > 1e-5 * dmax;
}

// one wave: sign-chosen LU of M (Q1 on entry, identity beyond nb) in place,
// the column rule of lila-c/lila_dorgh2_3.c:8-21: at step k, S_k = -1 when
// |1 - t| < |1 + t| for the diagonal t (column k negated), the pivot is
// S_k t - 1, the column below is divided by it and the trailing block gets
// the rank-1 update.  Lane c keeps column c in registers; the negation of
// column k is carried as a sign (the steps before k are linear in column
// k) and applied to its U entries at the end.  U (upper, diagonal the
// pivots) and L (strict lower) overwrite M, S gets the signs and D the
// reciprocal pivots.  Returns false when some |t| > 1 - 1e-10 (tau within
// 1e-10 of 0 or 2).
__device__ __forceinline__ bool wave_lu_sign32(double *M, double *S, double *D, int nb) {
  const int lane = opaque_lane(), c = lane & 31;
  double a[PNB];
#pragma unroll
  for (int r = 0; r < PNB; r++) a[r] = M[r * LDM + c];
  bool ok = true;
  double mys = 1.0, myinv = 1.0;
#pragma unroll
  for (int k = 0; k < PNB; k++) {
    if (k < nb) {
      const double t = rdlane(a[k], k);
      const bool flip = fabs(1.0 - t) < fabs(1.0 + t);
      if (fabs(t) > 1.0 - 1e-10) ok = false;
      const double sk = flip ? -1.0 : 1.0;
      const double piv = sk * t - 1.0;
      const double pinv = 1.0 / piv;
      const double f = sk * pinv;
      // U(k,c) for the trailing columns; finished columns c < k keep theirs
      const double uk = (c > k) ? a[k] : 0.0;
#pragma unroll
      for (int r = k + 1; r < PNB; r++) {
        const double l = rdlane(a[r], k) * f;  // L(r,k)
        a[r] = (c == k) ? l : a[r] - l * uk;
      }
      if (c == k) {
        a[k] = piv;
        mys = sk;
        myinv = pinv;
      }
    }
  }
  if (lane < PNB) {
#pragma unroll
    for (int r = 0; r < PNB; r++) M[r * LDM + c] = (r < c) ? a[r] * mys : a[r];
    S[c] = mys;
    D[c] = myinv;
  }
  wfence();
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

}  // namespace small32

#endif
