// linalg.hip -- triangular solves and products, SYRK, Cholesky, the
// dlarft T factor and the sign-chosen LU of Householder reconstruction,
// all on top of the FP64 MFMA GEMM (gemm_f64.hip).
//
// TRSM is recursive: the triangle is split at a multiple of the diagonal
// block (64), the off-diagonal part becomes one large GEMM, and each 64 x 64
// leaf is applied as a GEMM with its explicitly inverted diagonal block
// (all leaves inverted up front in one launch, one workgroup per block).
// Cholesky is recursive in the same way (TRSM + SYRK-form GEMM), with a
// single-workgroup factorization of each 64 x 64 leaf (the one-wave 32 x 32
// routines of small32.h, the column loop when a pivot fails).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "lila_dev.h"
#include "small32.h"

extern "C" int *ld_device_flags(void);

namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));
using small32::LDM;
constexpr int B32 = 32 * LDM;  // one 32 x 32 block (row-major, ld LDM)
constexpr int WG = 256;        // threads of the 64 x 64 leaf kernels

// block (bi, bj), bi <= bj, of an upper 64 x 64 matrix stored as 3 blocks
__device__ __forceinline__ double *ub(double *M, int bi, int bj) { return M + (bi + bj) * B32; }
__device__ __forceinline__ double &uel(double *M, int r, int c) {
  return M[((r >> 5) + (c >> 5)) * B32 + (r & 31) * LDM + (c & 31)];
}

// X = M^-1 for the upper 64 x 64 M in 3-block storage (lower parts zero),
// Dg = 1 / diag(M): the two diagonal 32 x 32 inverses by waves 0 and 1 at
// once, then X12 = -(X11 M12) X22 on MFMA (Tm: one free block)
__device__ void inv_upper64(double *M, const double *Dg, double *X, double *Tm) {
  const int wave = threadIdx.x >> 6;
  if (wave == 0) small32::inv_upper32(ub(M, 0, 0), Dg, ub(X, 0, 0));
  else if (wave == 1) small32::inv_upper32(ub(M, 1, 1), Dg + 32, ub(X, 1, 1));
  __syncthreads();
  small32::wg_mm32(ub(X, 0, 0), LDM, 1, false, ub(M, 0, 1), LDM, 1, Tm, 1.0, false);
  small32::wg_mm32(Tm, LDM, 1, false, ub(X, 1, 1), LDM, 1, ub(X, 0, 1), -1.0, false);
}

constexpr int TB = 64;  // diagonal block of TRSM / Cholesky leaves

inline hipStream_t S() { return (hipStream_t)ld_stream(); }

inline void gemm(int ta, int tb, int m, int n, int k, double alpha, const double *A, int lda,
                 int sa, const double *B, int ldb, int sb, double beta, double *C, int ldc,
                 int c_upper = 0) {
  ld_gemm_t g;
  memset(&g, 0, sizeof g);
  g.m = m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = lda; g.ta = ta; g.sa = sa;
  g.B = B; g.ldb = ldb; g.tb = tb; g.sb = sb;
  g.C = C; g.ldc = ldc; g.c_upper = c_upper;
  ld_gemm(&g);
}

// inverse of each TB x TB diagonal block of the stored triangle (uplo,
// unit) -> D[b] (TB x TB, zero outside the triangle), one workgroup of 256
// threads per block on the one-wave 32 x 32 routines and MFMA (a lower
// triangle is inverted as its transpose)
__global__ __launch_bounds__(WG) void trtri64_kernel(int nn, int upper, int unit,
                                                     const double *__restrict__ A, int lda,
                                                     double *__restrict__ D) {
  __shared__ double M[3 * B32], X[3 * B32], Tm[B32], Dg[64];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int o = b * TB, bs = min(TB, nn - o);
  // a thread's 16 values (row tid & 63, every fourth column) loaded into
  // registers first, so that the loads are in flight together (a load-store
  // loop waits out one memory latency per element)
  double v[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = tid & 63, c = (tid >> 6) + 4 * i;  // M(r, c), upper
    double x = 0.0;
    if (r < bs && c < bs) {
      if (r < c) x = upper ? A[(size_t)(o + r) + (size_t)(o + c) * lda] : A[(size_t)(o + c) + (size_t)(o + r) * lda];
      else if (r == c) x = unit ? 1.0 : A[(size_t)(o + r) + (size_t)(o + r) * lda];
    } else if (r == c) {
      x = 1.0;
    }
    v[i] = x;
  }
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = tid & 63, c = (tid >> 6) + 4 * i;
    if (!(r >= 32 && c < 32)) uel(M, r, c) = v[i];
  }
  __syncthreads();
  if (tid < 64) Dg[tid] = 1.0 / uel(M, tid, tid);
  __syncthreads();
  inv_upper64(M, Dg, X, Tm);
  double *Db = D + (size_t)b * TB * TB;
  for (int t = tid; t < 64 * 64; t += WG) {
    const int r = t & 63, c = t >> 6;  // D(r, c), column-major
    double v = 0.0;
    if (r < bs && c < bs) {
      if (upper && r <= c) v = uel(X, r, c);
      else if (!upper && r >= c) v = uel(X, c, r);
    }
    Db[r + (size_t)c * TB] = v;
  }
}

__global__ void scale_kernel(int m, int n, double alpha, double *B, int ldb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i < m && j < n) B[(size_t)i + (size_t)j * ldb] *= alpha;
}

struct TrsmCtx {
  char side;
  bool lowerop;  // op(A) lower triangular
  int tA;        // op = transpose
  const double *A;
  int lda;
  const double *D;  // inverted diagonal blocks
  double *B;
  int ldb;
  int m, n;  // B is m x n
};

// op(A)[r0:, c0:] as a GEMM operand (pointer; transpose flag is tA)
inline const double *opA(const TrsmCtx &t, int r0, int c0) {
  return t.tA ? t.A + c0 + (size_t)r0 * t.lda : t.A + r0 + (size_t)c0 * t.lda;
}

// a leaf of the blocked TRSM whose B is narrow, in one launch and in place:
// B_o <- op(D) B_o (LEFT: a workgroup per 32 columns of B, all len rows) or
// B_o <- B_o op(D) (right: a workgroup per 64 rows, all len columns), D the
// leaf's inverted diagonal block (TB x TB, zero outside its triangle).  A
// workgroup reads everything it writes before its first store, and no two
// workgroups share an element, so the product needs no copy of B (the GEMM
// form writes a workspace and copies it back: two launches).  Sums run in
// the order of k, one FMA per term.
constexpr int TL_LD = TB + 1;   // LDS row stride
constexpr int TL_MAX = 2048;    // the narrow side's largest extent taken here
template <bool LEFT>
__global__ __launch_bounds__(WG) void trsm_leaf_kernel(int len, int other, int tA, const double *__restrict__ D,
                                                       double *B, int ldb) {
  __shared__ double Ds[TB * TL_LD], Bs[TB * TL_LD];
  const int tid = threadIdx.x;
  // op(D) into LDS: LEFT as Ds[r][k] = op(D)(r, k), right as Ds[k][c] =
  // op(D)(k, c); every load of a thread in flight together
  {
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int q = tid + i * WG;
      v[i] = D[q];  // D(q & 63, q >> 6), column-major, ld TB
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int q = tid + i * WG, di = q & 63, dj = q >> 6;
      // LEFT: op(D)(r, k) = tA ? D(k, r) : D(r, k); right: op(D)(k, c) = tA ? D(c, k) : D(k, c)
      const int row = tA ? dj : di, col = tA ? di : dj;
      Ds[row * TL_LD + col] = v[i];
    }
  }
  if (LEFT) {
    const int c0 = blockIdx.x * 32;
    double v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int q = tid + i * WG, k = q & 63, c = q >> 6;
      v[i] = (k < len && c0 + c < other) ? B[(size_t)k + (size_t)(c0 + c) * ldb] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int q = tid + i * WG, k = q & 63, c = q >> 6;
      Bs[k * TL_LD + c] = v[i];
    }
    __syncthreads();
    const int c = tid & 31, rg = tid >> 5;
    double acc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) acc[i] = 0.0;
    for (int k = 0; k < len; k++) {
      const double b = Bs[k * TL_LD + c];
#pragma unroll
      for (int i = 0; i < 8; i++) acc[i] = fma(Ds[(rg + 8 * i) * TL_LD + k], b, acc[i]);
    }
    if (c0 + c < other)
#pragma unroll
      for (int i = 0; i < 8; i++)
        if (rg + 8 * i < len) B[(size_t)(rg + 8 * i) + (size_t)(c0 + c) * ldb] = acc[i];
  } else {
    const int r0 = blockIdx.x * 64;
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int q = tid + i * WG, r = q & 63, k = q >> 6;
      v[i] = (k < len && r0 + r < other) ? B[(size_t)(r0 + r) + (size_t)k * ldb] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int q = tid + i * WG, r = q & 63, k = q >> 6;
      Bs[k * TL_LD + r] = v[i];
    }
    __syncthreads();
    const int r = tid & 63, cg = tid >> 6;
    double acc[16];
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0;
    for (int k = 0; k < len; k++) {
      const double b = Bs[k * TL_LD + r];
#pragma unroll
      for (int i = 0; i < 16; i++) acc[i] = fma(b, Ds[k * TL_LD + cg + 4 * i], acc[i]);
    }
    if (r0 + r < other)
#pragma unroll
      for (int i = 0; i < 16; i++)
        if (cg + 4 * i < len) B[(size_t)(r0 + r) + (size_t)(cg + 4 * i) * ldb] = acc[i];
  }
}

// solve on the index range [o, o+len) of the triangle (multiple of TB at o)
void trsm_rec(const TrsmCtx &t, int o, int len) {
  if (len <= TB) {
    const double *Db = t.D + (size_t)(o / TB) * TB * TB;
    if ((t.side == 'L' ? t.n : t.m) <= TL_MAX) {
      if (t.side == 'L')
        hipLaunchKernelGGL(trsm_leaf_kernel<true>, dim3((t.n + 31) / 32), dim3(WG), 0, S(), len, t.n, t.tA, Db,
                           t.B + o, t.ldb);
      else
        hipLaunchKernelGGL(trsm_leaf_kernel<false>, dim3((t.m + 63) / 64), dim3(WG), 0, S(), len, t.m, t.tA, Db,
                           t.B + (size_t)o * t.ldb, t.ldb);
      return;
    }
    size_t mark = ld_ws_mark();
    if (t.side == 'L') {
      double *W = (double *)ld_ws_alloc(sizeof(double) * (size_t)len * t.n);
      gemm(t.tA, 0, len, t.n, len, 1.0, Db, TB, LD_GEN, t.B + o, t.ldb, LD_GEN, 0.0, W, len);
      ld_lacpy('A', len, t.n, W, len, t.B + o, t.ldb);
    } else {
      double *W = (double *)ld_ws_alloc(sizeof(double) * (size_t)t.m * len);
      gemm(0, t.tA, t.m, len, len, 1.0, t.B + (size_t)o * t.ldb, t.ldb, LD_GEN, Db, TB, LD_GEN,
           0.0, W, t.m);
      ld_lacpy('A', t.m, len, W, t.m, t.B + (size_t)o * t.ldb, t.ldb);
    }
    ld_ws_release(mark);
    return;
  }
  int nblk = (len + TB - 1) / TB;
  int l1 = ((nblk + 1) / 2) * TB;
  int l2 = len - l1;
  int o1 = o, o2 = o + l1;
  if (t.side == 'L') {
    if (t.lowerop) {
      trsm_rec(t, o1, l1);
      // B2 -= op(A)[o2, o1] X1
      gemm(t.tA, 0, l2, t.n, l1, -1.0, opA(t, o2, o1), t.lda, LD_GEN, t.B + o1, t.ldb, LD_GEN,
           1.0, t.B + o2, t.ldb);
      trsm_rec(t, o2, l2);
    } else {
      trsm_rec(t, o2, l2);
      gemm(t.tA, 0, l1, t.n, l2, -1.0, opA(t, o1, o2), t.lda, LD_GEN, t.B + o2, t.ldb, LD_GEN,
           1.0, t.B + o1, t.ldb);
      trsm_rec(t, o1, l1);
    }
  } else {
    if (!t.lowerop) {
      trsm_rec(t, o1, l1);
      // B2 -= X1 op(A)[o1, o2]
      gemm(0, t.tA, t.m, l2, l1, -1.0, t.B + (size_t)o1 * t.ldb, t.ldb, LD_GEN, opA(t, o1, o2),
           t.lda, LD_GEN, 1.0, t.B + (size_t)o2 * t.ldb, t.ldb);
      trsm_rec(t, o2, l2);
    } else {
      trsm_rec(t, o2, l2);
      gemm(0, t.tA, t.m, l1, l2, -1.0, t.B + (size_t)o2 * t.ldb, t.ldb, LD_GEN, opA(t, o2, o1),
           t.lda, LD_GEN, 1.0, t.B + (size_t)o1 * t.ldb, t.ldb);
      trsm_rec(t, o1, l1);
    }
  }
}

// unblocked upper Cholesky of one leaf: U = chol(A_bb(upper) - Sub), Sub
// optional (bs x bs, ld bs); flags[1] gets the first failing global column.
__device__ void potf2_body(int bs, double *A, int lda, const double *Sub, int col0, int *flags,
                           double (&U)[TB][TB + 1], int &bad);

// upper Cholesky of a leaf of at most 64 columns by 256 threads: R11 =
// chol(A11) and R22 = chol(A22 - R12^T R12) by the one-wave 32 x 32 routine
// (waves 1-3 idle meanwhile), R12 = R11^-T A12 on MFMA.  A pivot that is not
// positive sends the leaf to the column-by-column loop (threads 0..63), which
// also finds the first failing column for the LAPACK-style info.
__global__ __launch_bounds__(WG) void potf2_64_kernel(int bs, double *A, int lda, int col0, int *flags) {
  __shared__ double M[3 * B32], Sc[B32], Dg[64];
  __shared__ int okf;
  const int tid = threadIdx.x, wave = tid >> 6;
  {
    // all of a thread's loads in flight together (as in trtri64_kernel)
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int r = tid & 63, c = (tid >> 6) + 4 * i;
      v[i] = (r < bs && c < bs) ? (r <= c ? A[(size_t)r + (size_t)c * lda] : 0.0) : (r == c ? 1.0 : 0.0);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int r = tid & 63, c = (tid >> 6) + 4 * i;
      if (!(r >= 32 && c < 32)) uel(M, r, c) = v[i];
    }
  }
  __syncthreads();
  if (wave == 0) {
    const bool ok = small32::wave_chol32(ub(M, 0, 0), Dg, bs < 32 ? bs : 32, false);
    if (tid == 0) okf = ok ? 1 : 0;
  }
  __syncthreads();
  if (okf) {
    if (wave == 0) small32::inv_upper32(ub(M, 0, 0), Dg, Sc);
    __syncthreads();
    // R12 = R11^-T A12, then A22 -= R12^T R12
    small32::wg_mm32(Sc, 1, LDM, false, ub(M, 0, 1), LDM, 1, ub(M, 0, 1), 1.0, false);
    small32::wg_mm32(ub(M, 0, 1), 1, LDM, false, ub(M, 0, 1), LDM, 1, ub(M, 1, 1), -1.0, true);
    if (wave == 0) {
      const bool ok = small32::wave_chol32(ub(M, 1, 1), Dg + 32, bs > 32 ? bs - 32 : 0, false);
      if (tid == 0) okf = ok ? 1 : 0;
    }
    __syncthreads();
  }
  if (okf) {
    for (int t = tid; t < 64 * 64; t += WG) {
      const int r = t & 63, c = t >> 6;
      if (r <= c && c < bs) A[(size_t)r + (size_t)c * lda] = uel(M, r, c);
    }
    return;
  }
  // the column loop on the leaf as it came in (A is still unchanged); all
  // threads take part in its barriers, threads 0..63 do the work
  __shared__ double U[TB][TB + 1];
  __shared__ int bad;
  potf2_body(bs, A, lda, nullptr, col0, flags, U, bad);
}

__device__ void potf2_body(int bs, double *A, int lda, const double *Sub, int col0, int *flags,
                           double (&U)[TB][TB + 1], int &bad) {
  // thread c < TB keeps column c; threads past TB (a larger workgroup) only
  // take part in the barriers
  const int c = threadIdx.x;
  const bool mine = c < TB;
  if (c == 0) bad = 0;
  if (mine)
    for (int r = 0; r < TB; r++) {
      double v = 0.0;
      if (r < bs && c < bs && r <= c) {
        v = A[(size_t)r + (size_t)c * lda];
        if (Sub) v -= Sub[r + (size_t)c * bs];
      }
      U[r][c] = v;
    }
  __syncthreads();
  for (int j = 0; j < bs; j++) {
    // U(j,j) and row j of U
    if (c == 0) {
      double d = U[j][j];
      if (!(d > 0.0) && !bad) {
        bad = 1;
        atomicCAS(&flags[1], 0, col0 + j + 1);
      }
      U[j][j] = (d > 0.0) ? sqrt(d) : d;
    }
    __syncthreads();
    if (mine && c > j && c < bs) U[j][c] /= U[j][j];
    __syncthreads();
    if (mine && c > j && c < bs)
      for (int r = j + 1; r <= c; r++) U[r][c] -= U[j][r] * U[j][c];
    __syncthreads();
  }
  if (mine)
    for (int r = 0; r < bs; r++)
      if (c < bs && r <= c) A[(size_t)r + (size_t)c * lda] = U[r][c];
}

void potrf_rec(int n, double *A, int lda, int col0, const double *Dscratch_unused) {
  (void)Dscratch_unused;
  if (n <= TB) {
    hipLaunchKernelGGL(potf2_64_kernel, dim3(1), dim3(WG), 0, S(), n, A, lda, col0, ld_device_flags());
    return;
  }
  int nblk = (n + TB - 1) / TB;
  int n1 = ((nblk + 1) / 2) * TB, n2 = n - n1;
  potrf_rec(n1, A, lda, col0, nullptr);
  // A12 = U11^{-T} A12
  ld_trsm('L', 'U', 'T', 'N', n1, n2, 1.0, A, lda, A + (size_t)n1 * lda, lda);
  // A22 -= A12^T A12 (upper part only)
  gemm(1, 0, n2, n2, n1, -1.0, A + (size_t)n1 * lda, lda, LD_GEN, A + (size_t)n1 * lda, lda, LD_GEN,
       1.0, A + n1 + (size_t)n1 * lda, lda, 1);
  potrf_rec(n2, A + n1 + (size_t)n1 * lda, lda, col0 + n1, nullptr);
}

// small dlarft recurrence from the Gram G = V^T V (k <= 32, one workgroup)
__global__ void larft_small_kernel(int k, const double *__restrict__ Gm, int ldg,
                                   const double *__restrict__ tau, double *T, int ldt) {
  __shared__ double Tl[32][33];
  __shared__ double tmp[32];
  const int t = threadIdx.x;
  for (int r = 0; r < 32; r++) Tl[r][t] = 0.0;
  __syncthreads();
  for (int j = 0; j < k; j++) {
    const double tj = tau[j];
    if (t < j) tmp[t] = -tj * Gm[t + (size_t)j * ldg];
    __syncthreads();
    if (t < j) {
      double s = 0.0;
      if (tj != 0.0)
        for (int q = t; q < j; q++) s += Tl[t][q] * tmp[q];
      Tl[t][j] = s;
    }
    if (t == j) Tl[j][j] = tj;
    __syncthreads();
  }
  for (int r = 0; r < k; r++)
    if (t < k && r <= t) T[(size_t)r + (size_t)t * ldt] = Tl[r][t];
}

// sign-chosen unpivoted LU of an n x n block (one workgroup).
// rule 0 (dorgh2): if |1 - t_kk| < |1 + t_kk| negate column k, S = -1;
//                  then t_kk -= 1.
// rule 1 (dgetrf_b): if |1 - a_kk| < |1 + a_kk| keep the column, S = +1,
//                  else negate the column, S = -1; then a_kk += 1.
// Column negation covers rows [0, n) of the block; scaling and the rank-1
// update cover rows [0, nrow) so a tall block factors in one pass.
__global__ void lu_sign_kernel(int n, int nrow, double *T, int ldt, double *Sg, int rule) {
  __shared__ double piv;
  __shared__ int neg;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int k = 0; k < n; k++) {
    if (tid == 0) {
      double t = T[(size_t)k + (size_t)k * ldt];
      bool c = fabs(1.0 - t) < fabs(1.0 + t);
      if (rule == 0) { neg = c ? 1 : 0; Sg[k] = c ? -1.0 : 1.0; }
      else { neg = c ? 0 : 1; Sg[k] = c ? 1.0 : -1.0; }
    }
    __syncthreads();
    if (neg)
      for (int r = tid; r < nrow; r += nt) T[(size_t)r + (size_t)k * ldt] = -T[(size_t)r + (size_t)k * ldt];
    __syncthreads();
    if (tid == 0) {
      double t = T[(size_t)k + (size_t)k * ldt] + (rule == 0 ? -1.0 : 1.0);
      T[(size_t)k + (size_t)k * ldt] = t;
      piv = t;
    }
    __syncthreads();
    for (int r = k + 1 + tid; r < nrow; r += nt) T[(size_t)r + (size_t)k * ldt] /= piv;
    __syncthreads();
    const int rows = nrow - k - 1, cols = n - k - 1;
    for (long e = tid; e < (long)rows * cols; e += nt) {
      int r = k + 1 + (int)(e % rows), c = k + 1 + (int)(e / rows);
      T[(size_t)r + (size_t)c * ldt] -= T[(size_t)r + (size_t)k * ldt] * T[(size_t)k + (size_t)c * ldt];
    }
    __syncthreads();
  }
}

// rule 0 on an mr x nc panel (nc <= 32) by 256 threads: the top 32 x 32
// (identity past the panel) by the one-wave sign-chosen LU -- the same
// pivots, signs and factors as the column loop above; its reciprocal pivots
// go to Dg for lu_rows_kernel, which solves the rows below
__global__ __launch_bounds__(WG) void lu_sign32_leaf_kernel(int nc, int mr, double *T, int ldt, double *Sg,
                                                            double *Dg) {
  __shared__ double M[B32], Sv[32], Dv[32];
  const int tid = threadIdx.x, wave = tid >> 6;
  {
    double v[4];  // all four loads of a thread in flight together
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int t = tid + i * WG, r = t & 31, c = t >> 5;
      v[i] = (r < mr && c < nc) ? T[(size_t)r + (size_t)c * ldt] : (r == c ? 1.0 : 0.0);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int t = tid + i * WG, r = t & 31, c = t >> 5;
      M[r * LDM + c] = v[i];
    }
  }
  __syncthreads();
  if (wave == 0) (void)small32::wave_lu_sign32(M, Sv, Dv, nc);
  __syncthreads();
  for (int t = tid; t < 32 * 32; t += WG) {
    const int r = t & 31, c = t >> 5;
    if (r < mr && c < nc) T[(size_t)r + (size_t)c * ldt] = M[r * LDM + c];
  }
  if (tid < nc) Sg[tid] = Sv[tid];
  if (tid < 32) Dg[tid] = Dv[tid];
}

// the rows below the top 32 of lu_sign32_leaf_kernel's panel, LU_RB rows a
// workgroup (a kernel of their own, the results stored through LDS: computed
// inline beside the one-wave LU and stored from registers, a row's values
// and their 32 addresses went to scratch memory, 60-140 us a leaf).  One row per
// thread, x U = a S by forward substitution against U (the leaf's output rows
// 0..31: strict upper part, signs applied) with the leaf's signs and
// reciprocal pivots, right-looking: x_c2 -= x_c U(c, c2) for c < c2 in the
// order of c, the same operations as the left-looking sum
constexpr int LU_RB = 128, LU_LDX = LU_RB + 1;
__global__ __launch_bounds__(LU_RB) void lu_rows_kernel(int nc, int mr, double *T, int ldt, const double *Sg,
                                                        const double *Dg) {
  __shared__ double U[32 * LDM], Sv[32], Dv[32], X[32 * LU_LDX];
  const int tid = threadIdx.x, r0 = 32 + blockIdx.x * LU_RB, nr = min(LU_RB, mr - r0);
  // every load of a thread issued before the first is used: U's strict upper
  // triangle (8 values a thread) and the thread's row (32 values; a wave
  // reads 64 consecutive rows of a column at a time)
  double u[8], x[32];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int t = tid + i * LU_RB, r = t & 31, c = t >> 5;
    u[i] = (r < c && c < nc) ? T[(size_t)r + (size_t)c * ldt] : 0.0;
  }
  const bool row = tid < nr;
  const double *xr = T + (size_t)(r0 + (row ? tid : 0));
#pragma unroll
  for (int c = 0; c < 32; c++) x[c] = (row && c < nc) ? xr[(size_t)c * ldt] : 0.0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int t = tid + i * LU_RB, r = t & 31, c = t >> 5;
    U[r * LDM + c] = u[i];
  }
  if (tid < 32) {
    Sv[tid] = (tid < nc) ? Sg[tid] : 1.0;
    Dv[tid] = (tid < nc) ? Dg[tid] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 32; c++) x[c] *= Sv[c];
#pragma unroll
  for (int c = 0; c < 32; c++) {
    x[c] *= Dv[c];
#pragma unroll
    for (int c2 = c + 1; c2 < 32; c2++) x[c2] -= x[c] * U[c * LDM + c2];
  }
  // through LDS to the stores, so that no address of the loads stays live
#pragma unroll
  for (int c = 0; c < 32; c++) X[c * LU_LDX + tid] = x[c];
  __syncthreads();
  for (int t = tid; t < 32 * LU_RB; t += LU_RB) {
    const int rl = t % LU_RB, c = t / LU_RB;
    if (rl < nr && c < nc) T[(size_t)(r0 + rl) + (size_t)c * ldt] = X[c * LU_LDX + rl];
  }
}

__global__ void flip_where_kernel(int m, int n, double *X, int ldx, const double *S, double which) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i < m && j < n && S[j] == which) X[(size_t)i + (size_t)j * ldx] = -X[(size_t)i + (size_t)j * ldx];
}

__global__ void flip_urows_where_kernel(int n, double *U, int ldu, const double *S, double which) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int c = blockIdx.y;
  if (k < n && c < n && c >= k && S[k] == which) U[(size_t)k + (size_t)c * ldu] = -U[(size_t)k + (size_t)c * ldu];
}


// Column j's share of ||I - G||_F^2 for the symmetric G whose upper triangle
// is stored: 2 sum_{i<j} G_ij^2 + (1 - G_jj)^2.  One workgroup of 256 per
// column, strided partials reduced in a fixed tree, so the sum does not
// depend on timing.
__global__ void orth_defect_cols_kernel(int n, const double *__restrict__ G, int ldg,
                                        double *__restrict__ part) {
  __shared__ double red[256];
  const int j = blockIdx.x, t = threadIdx.x;
  const double *c = G + (size_t)j * ldg;
  double s = 0.0;
  for (int i = t; i < j; i += 256) s += c[i] * c[i];
  s *= 2.0;
  if (t == 0) {
    const double d = 1.0 - c[j];
    s += d * d;
  }
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) part[j] = red[0];
}

// F = I - G in full, for the symmetric G whose upper triangle is stored
__global__ void sym_defect_full_kernel(int n, const double *__restrict__ G, int ldg,
                                       double *__restrict__ F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  const double g = (i <= j) ? G[i + (size_t)j * ldg] : G[j + (size_t)i * ldg];
  F[i + (size_t)j * n] = (i == j ? 1.0 : 0.0) - g;
}

// a fixed start vector in [-1, 1] (splitmix64 of the index)
__global__ void start_vector_kernel(int n, double *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  x[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// part[b] = sum of y[i]^2 over the b-th 256-element slice
__global__ void sumsq_slices_kernel(int n, const double *__restrict__ y, double *__restrict__ part) {
  __shared__ double red[256];
  const int t = threadIdx.x, i = blockIdx.x * 256 + t;
  const double v = i < n ? y[i] : 0.0;
  red[t] = v * v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = red[0];
}

// x = y / s[0] (s[0] > 0; zero vector when s[0] == 0)
__global__ void scale_by_kernel(int n, const double *__restrict__ y, const double *__restrict__ s,
                                double *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = s[0] > 0.0 ? y[i] / s[0] : 0.0;
}

// out[0] = sqrt(sum of part[0..n)), one workgroup, fixed order
__global__ void sum_sqrt_kernel(int n, const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int i = t; i < n; i += 256) s += part[i];
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) out[0] = sqrt(red[0]);
}

}  // namespace

extern "C" {

void ld_trsm(char side, char uplo, char trans, char diag, int m, int n, double alpha,
             const double *A, int lda, double *B, int ldb) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (m <= 0 || n <= 0) return;
  if (alpha != 1.0)
    hipLaunchKernelGGL(scale_kernel, dim3((m + 255) / 256, n), dim3(256), 0, S(), m, n, alpha, B, ldb);
  if (alpha == 0.0) return;
  const int nn = (side == 'L') ? m : n;
  const int nblk = (nn + TB - 1) / TB;
  size_t mark = ld_ws_mark();
  double *D = (double *)ld_ws_alloc(sizeof(double) * (size_t)nblk * TB * TB);
  hipLaunchKernelGGL(trtri64_kernel, dim3(nblk), dim3(WG), 0, S(), nn, uplo == 'U' ? 1 : 0,
                     diag == 'U' ? 1 : 0, A, lda, D);
  TrsmCtx t;
  t.side = side;
  t.tA = (trans == 'T') ? 1 : 0;
  t.lowerop = ((uplo == 'L') != (t.tA == 1));
  t.A = A; t.lda = lda; t.D = D; t.B = B; t.ldb = ldb; t.m = m; t.n = n;
  trsm_rec(t, 0, nn);
  ld_ws_release(mark);
}

void ld_trmm(char side, char uplo, char trans, char diag, int m, int n, double alpha,
             const double *A, int lda, double *B, int ldb) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (m <= 0 || n <= 0) return;
  int s = (uplo == 'U') ? (diag == 'U' ? LD_UUNIT : LD_UPPER) : (diag == 'U' ? LD_VUNIT : LD_LOWER);
  int tA = (trans == 'T') ? 1 : 0;
  size_t mark = ld_ws_mark();
  double *W = (double *)ld_ws_alloc(sizeof(double) * (size_t)m * n);
  if (side == 'L') gemm(tA, 0, m, n, m, alpha, A, lda, s, B, ldb, LD_GEN, 0.0, W, m);
  else gemm(0, tA, m, n, n, alpha, B, ldb, LD_GEN, A, lda, s, 0.0, W, m);
  ld_lacpy('A', m, n, W, m, B, ldb);
  ld_ws_release(mark);
}

void ld_syrk_ut(int n, int k, double alpha, const double *A, int lda, double beta, double *C, int ldc) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (n <= 0) return;
  gemm(1, 0, n, n, k, alpha, A, lda, LD_GEN, A, lda, LD_GEN, beta, C, ldc, 1);
}

void ld_potrf_u(int n, double *A, int lda) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (n <= 0) return;
  potrf_rec(n, A, lda, 0, nullptr);
}

void ld_larft(int ml, int k, const double *V, int ldv, const double *tau, double *T, int ldt) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (k <= 0) return;
  if (k <= 32) {
    size_t mark = ld_ws_mark();
    double *Gm = (double *)ld_ws_alloc(sizeof(double) * (size_t)k * k);
    gemm(1, 0, k, k, ml, 1.0, V, ldv, LD_VUNIT, V, ldv, LD_VUNIT, 0.0, Gm, k);
    hipLaunchKernelGGL(larft_small_kernel, dim3(1), dim3(32), 0, S(), k, Gm, k, tau, T, ldt);
    ld_ws_release(mark);
    return;
  }
  int k1 = k / 2, k2 = k - k1;
  ld_larft(ml, k1, V, ldv, tau, T, ldt);
  ld_larft(ml - k1, k2, V + k1 + (size_t)k1 * ldv, ldv, tau + k1, T + k1 + (size_t)k1 * ldt, ldt);
  // T12 = -T11 (V1^T V2) T22, V1^T V2 over rows k1.. (V2 is zero above)
  size_t mark = ld_ws_mark();
  double *G12 = (double *)ld_ws_alloc(sizeof(double) * (size_t)k1 * k2);
  double *H = (double *)ld_ws_alloc(sizeof(double) * (size_t)k1 * k2);
  gemm(1, 0, k1, k2, ml - k1, 1.0, V + k1, ldv, LD_GEN, V + k1 + (size_t)k1 * ldv, ldv, LD_VUNIT, 0.0,
       G12, k1);
  gemm(0, 0, k1, k2, k1, 1.0, T, ldt, LD_UPPER, G12, k1, LD_GEN, 0.0, H, k1);
  gemm(0, 0, k1, k2, k2, -1.0, H, k1, LD_GEN, T + k1 + (size_t)k1 * ldt, ldt, LD_UPPER, 0.0,
       T + (size_t)k1 * ldt, ldt);
  ld_ws_release(mark);
}

void ld_n2t(int n, const double *tau, double *T, int ldt) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (n <= 0) return;
  if (n <= 32) {
    size_t mark = ld_ws_mark();
    double *Nm = (double *)ld_ws_alloc(sizeof(double) * (size_t)n * n);
    ld_lacpy('U', n, n, T, ldt, Nm, n);
    hipLaunchKernelGGL(larft_small_kernel, dim3(1), dim3(32), 0, S(), n, Nm, n, tau, T, ldt);
    ld_ws_release(mark);
    return;
  }
  const int n1 = n / 2, n2 = n - n1;
  double *T22 = T + n1 + (size_t)n1 * ldt, *T12 = T + (size_t)n1 * ldt;
  ld_n2t(n1, tau, T, ldt);
  ld_n2t(n2, tau + n1, T22, ldt);
  // T12 = -T11 N12 T22
  size_t mark = ld_ws_mark();
  double *H = (double *)ld_ws_alloc(sizeof(double) * (size_t)n1 * n2);
  gemm(0, 0, n1, n2, n1, 1.0, T, ldt, LD_UPPER, T12, ldt, LD_GEN, 0.0, H, n1);
  gemm(0, 0, n1, n2, n2, -1.0, H, n1, LD_GEN, T22, ldt, LD_UPPER, 0.0, T12, ldt);
  ld_ws_release(mark);
}

// Sign-chosen unpivoted LU of the mr x nc panel A (mr >= nc), recursive over
// the columns: the left half, U12 = L11^-1 A12, A22 -= L21 U12 (FP64 MFMA),
// the right half; column k's negation (S_k = -1) reaches its U12 entries
// afterwards -- every step before k is linear in column k, so negating the
// column when it is reached or at the start gives the same factors.  Panels
// of at most 32 columns run the one-workgroup kernel.
static void lu_sign_panel(int mr, int nc, double *A, int lda, double *Sg, int rule) {
  if (nc <= 0) return;
  if (nc <= 32) {
    if (rule == 0) {
      // the leaf's reciprocal pivots pass to the rows kernel through Dg
      size_t mark = ld_ws_mark();
      double *Dg = (double *)ld_ws_alloc(32 * sizeof(double));
      if (!ld_failed()) {
        hipLaunchKernelGGL(lu_sign32_leaf_kernel, dim3(1), dim3(WG), 0, S(), nc, mr, A, lda, Sg, Dg);
        if (mr > 32)
          hipLaunchKernelGGL(lu_rows_kernel, dim3((mr - 32 + LU_RB - 1) / LU_RB), dim3(LU_RB), 0, S(), nc, mr, A,
                             lda, Sg, Dg);
      }
      ld_ws_release(mark);
    } else {
      hipLaunchKernelGGL(lu_sign_kernel, dim3(1), dim3(1024), 0, S(), nc, mr, A, lda, Sg, rule);
    }
    return;
  }
  const int n1 = (nc / 2 + 31) / 32 * 32, n2 = nc - n1;
  lu_sign_panel(mr, n1, A, lda, Sg, rule);
  double *A12 = A + (size_t)n1 * lda, *A21 = A + n1, *A22 = A + n1 + (size_t)n1 * lda;
  ld_trsm('L', 'L', 'N', 'U', n1, n2, 1.0, A, lda, A12, lda);
  ld_gemm_simple(0, 0, mr - n1, n2, n1, -1.0, A21, lda, LD_GEN, A12, lda, LD_GEN, 1.0, A22, lda);
  lu_sign_panel(mr - n1, n2, A22, lda, Sg + n1, rule);
  hipLaunchKernelGGL(flip_where_kernel, dim3((n1 + 255) / 256, n2), dim3(256), 0, S(), n1, n2, A12, lda,
                     Sg + n1, -1.0);
}

void ld_lu_sign(int n, double *Tq, int ldt, double *Sg) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (n <= 0) return;
  lu_sign_panel(n, n, Tq, ldt, Sg, 0);
}

// c3/lila_dgetrf_b.c semantics: A (ml x n) holds the orthonormal basis, Q
// (ml x n) is flipped where the column is kept, U (n x n, ld n) holds R and
// gets its rows flipped likewise.  Top block factored with the sign rule,
// rows below by L2 = A2 D U'^{-1}.
void ld_getrf_b(int ml, int n, double *A, int lda, double *Q, int ldq, double *U) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (n <= 0 || ml <= 0) return;
  size_t mark = ld_ws_mark();
  double *Sg = (double *)ld_ws_alloc(sizeof(double) * (size_t)n);
  lu_sign_panel(n, n, A, lda, Sg, 1);
  if (ml > n) {
    hipLaunchKernelGGL(flip_where_kernel, dim3((ml - n + 255) / 256, n), dim3(256), 0, S(), ml - n, n,
                       A + n, lda, Sg, -1.0);
    ld_trsm('R', 'U', 'N', 'N', ml - n, n, 1.0, A, lda, A + n, lda);
  }
  hipLaunchKernelGGL(flip_where_kernel, dim3((ml + 255) / 256, n), dim3(256), 0, S(), ml, n, Q, ldq,
                     Sg, 1.0);
  hipLaunchKernelGGL(flip_urows_where_kernel, dim3((n + 255) / 256, n), dim3(256), 0, S(), n, U, n,
                     Sg, 1.0);
  ld_ws_release(mark);
}


// ||I - G||_F for the symmetric n x n G whose upper triangle is stored; waits
// for the current lane's stream.  Used by the mixed-precision driver to
// decide whether a second Cholesky-QR pass is needed.
double ld_orth_defect(int n, const double *G, int ldg) {
  if (ld_failed()) return 0.0;  // an earlier allocation of this call failed
  if (n <= 0) return 0.0;
  size_t mark = ld_ws_mark();
  double *part = (double *)ld_ws_alloc(sizeof(double) * ((size_t)n + 1));
  hipLaunchKernelGGL(orth_defect_cols_kernel, dim3(n), dim3(256), 0, S(), n, G, ldg, part);
  hipLaunchKernelGGL(sum_sqrt_kernel, dim3(1), dim3(256), 0, S(), n, part, part + n);
  double d = NAN;
  if (hipMemcpyAsync(&d, part + n, sizeof d, hipMemcpyDeviceToHost, S()) != hipSuccess ||
      hipStreamSynchronize(S()) != hipSuccess)
    d = NAN;
  ld_ws_release(mark);
  return d;
}

// A lower estimate of ||I - G||_2 (G symmetric, upper triangle stored):
// `iters` steps of the power method on F = I - G from a fixed start vector,
// the last ||F x|| with ||x|| = 1.  Every reduction runs in a fixed order, so
// the estimate does not vary from run to run.  Waits for the stream.
double ld_sym_defect_norm2_est(int n, const double *G, int ldg, int iters) {
  if (ld_failed()) return 0.0;  // an earlier allocation of this call failed
  if (n <= 0) return 0.0;
  size_t mark = ld_ws_mark();
  const int nb = (n + 255) / 256;
  double *F = (double *)ld_ws_alloc(sizeof(double) * (size_t)n * n);
  double *x = (double *)ld_ws_alloc(sizeof(double) * (size_t)n);
  double *y = (double *)ld_ws_alloc(sizeof(double) * (size_t)n);
  double *part = (double *)ld_ws_alloc(sizeof(double) * ((size_t)nb + 1));
  hipLaunchKernelGGL(sym_defect_full_kernel, dim3(nb, n), dim3(256), 0, S(), n, G, ldg, F);
  hipLaunchKernelGGL(start_vector_kernel, dim3(nb), dim3(256), 0, S(), n, y);
  for (int k = 0; k <= iters; k++) {
    // x = y / ||y||, then (except after the last norm) y = F x
    hipLaunchKernelGGL(sumsq_slices_kernel, dim3(nb), dim3(256), 0, S(), n, y, part);
    hipLaunchKernelGGL(sum_sqrt_kernel, dim3(1), dim3(256), 0, S(), nb, part, part + nb);
    if (k == iters) break;
    hipLaunchKernelGGL(scale_by_kernel, dim3(nb), dim3(256), 0, S(), n, y, part + nb, x);
    gemm(0, 0, n, 1, n, 1.0, F, n, LD_GEN, x, n, LD_GEN, 0.0, y, n);
  }
  double d = NAN;
  if (hipMemcpyAsync(&d, part + nb, sizeof d, hipMemcpyDeviceToHost, S()) != hipSuccess ||
      hipStreamSynchronize(S()) != hipSuccess)
    d = NAN;
  ld_ws_release(mark);
  return d;
}

}  // extern "C"
