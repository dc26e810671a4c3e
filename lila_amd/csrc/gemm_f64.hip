// gemm_f64.hip -- FP64 MFMA GEMM for the LILA hot path (gfx950).
//
// C = alpha * op(A) * op(B) + beta * C, column-major, with per-operand
// structure masks so that Householder V (unit lower trapezoidal), T and R
// (upper triangular) are read in place without being copied out:
//   W = V^T C, W = T^T W, C -= V W (block-reflector application) and
//   G = V_j^T V_i (T-factor connection) are all single launches.
//
// Tiling: 256 threads = 4 waves in a 2 x 2 grid; block tile BM x BN with
// BK = 16; each wave owns (BM/2) x (BN/2) built from 16x16 tiles of
// v_mfma_f64_16x16x4_f64.  The MFMA operands are fed swapped (op(B)^T as the
// 16x4 "A" operand, op(A)^T as the 4x16 "B" operand) so the accumulator holds
// C^T fragments: lane l owns C row (l % 16) of columns (l/16) + 4v, which
// makes every store a run of 16 consecutive doubles of one column.
//
// Long-K shapes (M, N small, K = m - i in the hundreds of thousands) are
// split over grid.z into FP64 partials that a second kernel reduces in a
// fixed order (deterministic).
#include <hip/hip_runtime.h>
#include <type_traits>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include "lila_dev.h"
#include <rocblas/rocblas.h>

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// one MFMA step on either precision: FP64 v_mfma_f64_16x16x4f64, or FP32
// v_mfma_f32_16x16x4f32 (twice the rate) for the low-precision updates
__device__ __forceinline__ d4_t mfma16(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma16(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

namespace {

constexpr int BK = 16;
constexpr int NT = 256;

__device__ __forceinline__ double mask_val(const double *__restrict__ X, int ldx,
                                           int r, int c, int s) {
  // r, c: stored indices (already bounds-checked by the caller)
  if (s == LD_VUNIT) {
    if (r < c) return 0.0;
    if (r == c) return 1.0;
  } else if (s == LD_UPPER) {
    if (r > c) return 0.0;
  } else if (s == LD_LOWER) {
    if (r < c) return 0.0;
  } else if (s == LD_UUNIT) {
    if (r > c) return 0.0;
    if (r == c) return 1.0;
  }
  return X[(size_t)r + (size_t)c * ldx];
}

// does a K-tile of a structured stored operand touch the triangle's edge?
// (rows [r0, r1], cols [c0, c1] of the stored matrix, inclusive)
__device__ __forceinline__ bool tile_needs_mask(int s, int r0, int r1, int c0, int c1) {
  if (s == LD_GEN) return false;
  if (s == LD_VUNIT || s == LD_LOWER) return r0 <= c1;  // some r <= c
  return r1 >= c0;                                      // upper forms: some r >= c
}

// F32: operands rounded to FP32 when staged into LDS, FP32 MFMA with FP32
// accumulation, result widened to FP64 in the epilogue (C stays FP64)
template <int BM, int BN, bool TA, bool TB, bool F32 = false>
__global__ __launch_bounds__(NT, 2) void gemm_f64_kernel(
    int M, int N, int K, double alpha, double beta,
    const double *__restrict__ A, int lda, int sa,
    const double *__restrict__ B, int ldb, int sb,
    double *__restrict__ C, int ldc, int zero_rows, int c_upper,
    int kchunk, double *__restrict__ P) {
  constexpr int PA = 1, PB = 1;
  // SYRK form: a tile entirely below the diagonal writes nothing
  // XCD-aware tile order: workgroups are dispatched round-robin over the 8
  // XCDs (each with its own L2); give XCD x the contiguous x-th eighth of the
  // output tiles and walk it in 8 x 8 blocks, so the workgroups resident on
  // one XCD share A row panels and B column panels through that L2
  int tx = blockIdx.x, ty = blockIdx.y;
  {
    const int gx = gridDim.x, gy = gridDim.y, ntile = gx * gy;
    if (ntile >= 512 && gx % 8 == 0 && gy % 8 == 0) {
      const int lin = blockIdx.x + blockIdx.y * gx;
      const int t = (lin % 8) * (ntile / 8) + lin / 8;
      const int st = t / 64, w = t % 64, sx = gx / 8;
      tx = (st % sx) * 8 + w % 8;
      ty = (st / sx) * 8 + w / 8;
    }
  }
  if (c_upper && tx * BM > ty * BN + BN - 1) return;
  constexpr int WMT = BM / 32;  // 16x16 tiles per wave along m
  constexpr int WNT = BN / 32;  // along n
  constexpr int NLA = BM * BK / NT;
  constexpr int NLB = BN * BK / NT;
  // double-buffered tiles, k-major: As[buf][k][m], Bs[buf][k][n]
  typedef typename std::conditional<F32, float, double>::type lds_t;
  typedef typename std::conditional<F32, f4_t, d4_t>::type acc_t;
  __shared__ lds_t As[2][BK][BM + PA];
  __shared__ lds_t Bs[2][BK][BN + PB];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = tx * BM, n0 = ty * BN;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(K, kbeg + kchunk);

  // fixed per-thread load coordinates: element t of the A tile is
  // (i, p) = (ia, pa + t*da) or (ia + t*da, pa) depending on the layout
  // (NT is a multiple of BM, BN and BK)
  const int a_i0 = TA ? tid / BK : tid % BM;
  const int a_p0 = TA ? tid % BK : tid / BM;
  const int b_j0 = TB ? tid % BN : tid / BK;
  const int b_p0 = TB ? tid / BN : tid % BK;
  // pointer to stored element of op(A)(m0 + a_i0, a_p0) and its t-stride
  const double *Ab = TA ? A + a_p0 + (size_t)(m0 + a_i0) * lda : A + (m0 + a_i0) + (size_t)a_p0 * lda;
  const size_t a_ts = TA ? (size_t)(NT / BK) * lda : (size_t)(NT / BM) * lda;
  const size_t a_ks = TA ? 1 : (size_t)lda;  // stride of one k step
  const double *Bb = TB ? B + (n0 + b_j0) + (size_t)b_p0 * ldb : B + b_p0 + (size_t)(n0 + b_j0) * ldb;
  const size_t b_ts = TB ? (size_t)(NT / BN) * ldb : (size_t)(NT / BK) * ldb;
  const size_t b_ks = TB ? (size_t)ldb : 1;
  const bool a_rows_full = (m0 + BM <= M);
  const bool b_cols_full = (n0 + BN <= N);

  acc_t acc[WMT][WNT];
#pragma unroll
  for (int a = 0; a < WMT; a++)
#pragma unroll
    for (int b = 0; b < WNT; b++) acc[a][b] = acc_t{0, 0, 0, 0};

  double ra[NLA], rb[NLB];

  auto load_tiles = [&](int k0) {
    const bool kfull = (k0 + BK <= kend);
    const bool amask = TA ? tile_needs_mask(sa, k0, k0 + BK - 1, m0, m0 + BM - 1)
                          : tile_needs_mask(sa, m0, m0 + BM - 1, k0, k0 + BK - 1);
    const bool bmask = TB ? tile_needs_mask(sb, n0, n0 + BN - 1, k0, k0 + BK - 1)
                          : tile_needs_mask(sb, k0, k0 + BK - 1, n0, n0 + BN - 1);
    if (kfull && a_rows_full && !amask) {
      const double *p = Ab + (size_t)k0 * a_ks;
#pragma unroll
      for (int t = 0; t < NLA; t++) ra[t] = p[t * a_ts];
    } else {
#pragma unroll
      for (int t = 0; t < NLA; t++) {
        const int i = TA ? a_i0 + t * (NT / BK) : a_i0;
        const int pk = TA ? a_p0 : a_p0 + t * (NT / BM);
        const int gi = m0 + i, gp = k0 + pk;
        double v = 0.0;
        if (gi < M && gp < kend) v = TA ? mask_val(A, lda, gp, gi, sa) : mask_val(A, lda, gi, gp, sa);
        ra[t] = v;
      }
    }
    if (kfull && b_cols_full && !bmask) {
      const double *p = Bb + (size_t)k0 * b_ks;
#pragma unroll
      for (int t = 0; t < NLB; t++) rb[t] = p[t * b_ts];
    } else {
#pragma unroll
      for (int t = 0; t < NLB; t++) {
        const int j = TB ? b_j0 : b_j0 + t * (NT / BK);
        const int pk = TB ? b_p0 + t * (NT / BN) : b_p0;
        const int gj = n0 + j, gp = k0 + pk;
        double v = 0.0;
        if (gj < N && gp < kend) v = TB ? mask_val(B, ldb, gj, gp, sb) : mask_val(B, ldb, gp, gj, sb);
        rb[t] = v;
      }
    }
  };
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int t = 0; t < NLA; t++) {
      const int i = TA ? a_i0 + t * (NT / BK) : a_i0;
      const int pk = TA ? a_p0 : a_p0 + t * (NT / BM);
      As[buf][pk][i] = (lds_t)ra[t];
    }
#pragma unroll
    for (int t = 0; t < NLB; t++) {
      const int j = TB ? b_j0 : b_j0 + t * (NT / BK);
      const int pk = TB ? b_p0 + t * (NT / BN) : b_p0;
      Bs[buf][pk][j] = (lds_t)rb[t];
    }
  };

  // the k range this tile needs: where a structured op(A) (rows [m0,
  // m0 + BM)) or op(B) (columns [n0, n0 + BN)) is zero by its triangle, the
  // products are exact zeros, so those k steps are skipped (a product of two
  // triangles, R = R2 R1, costs a third of a full one).  The steps kept stay
  // on the kbeg grid; edge steps still mask element by element.
  int klo = kbeg, khi = kend;
  {
    const bool a_up = (sa == LD_UPPER || sa == LD_UUNIT), a_lo = (sa == LD_LOWER || sa == LD_VUNIT);
    const bool b_up = (sb == LD_UPPER || sb == LD_UUNIT), b_lo = (sb == LD_LOWER || sb == LD_VUNIT);
    if (TA ? a_lo : a_up) klo = max(klo, m0);       // op(A)(i, k) = 0 for k < i
    if (TA ? a_up : a_lo) khi = min(khi, m0 + BM);  // op(A)(i, k) = 0 for k > i
    if (TB ? b_lo : b_up) khi = min(khi, n0 + BN);  // op(B)(k, j) = 0 for k > j
    if (TB ? b_up : b_lo) klo = max(klo, n0);       // op(B)(k, j) = 0 for k < j
    klo = kbeg + ((klo - kbeg) / BK) * BK;
  }

  int buf = 0;
  if (klo < khi) {
    load_tiles(klo);
    store_tiles(0);
  }
  __syncthreads();
  for (int k0 = klo; k0 < khi; k0 += BK) {
    const bool more = (k0 + BK < khi);
    if (more) load_tiles(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK / 4; kk++) {
      const int kr = kk * 4 + (lane >> 4);
      lds_t af[WMT], bf[WNT];
#pragma unroll
      for (int a = 0; a < WMT; a++) af[a] = As[buf][kr][wm * (BM / 2) + a * 16 + (lane & 15)];
#pragma unroll
      for (int b = 0; b < WNT; b++) bf[b] = Bs[buf][kr][wn * (BN / 2) + b * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < WMT; a++)
#pragma unroll
        for (int b = 0; b < WNT; b++)
          acc[a][b] = mfma16(bf[b], af[a], acc[a][b]);
    }
    if (more) store_tiles(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // epilogue: the f64 16x16x4 accumulator puts D row (lane/16) + 4v in lane
  // `lane`, register v (rows interleaved by 4), D column lane%16; the f32
  // 16x16x4 one puts row 4 (lane/16) + v there (four consecutive rows).
  // D = C^T here, so lane holds C(row = base + lane%16, col = base + D row)
  // Every C load of the tile is issued before the first store, so the tile
  // pays one memory round trip instead of one per element.
  const int rl = lane & 15, cg = lane >> 4;
  const bool read_c = (P == nullptr) && beta != 0.0;
  double cv[WMT][WNT][4];
#pragma unroll
  for (int a = 0; a < WMT; a++) {
    const int gi = m0 + wm * (BM / 2) + a * 16 + rl;
#pragma unroll
    for (int b = 0; b < WNT; b++) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int gj = n0 + wn * (BN / 2) + b * 16 + (F32 ? 4 * cg + v : cg + 4 * v);
        const bool ld = read_c && gi < M && gj < N && gi >= zero_rows && !(c_upper && gi > gj);
        cv[a][b][v] = ld ? C[(size_t)gi + (size_t)gj * ldc] : 0.0;
      }
    }
  }
#pragma unroll
  for (int a = 0; a < WMT; a++) {
    const int gi = m0 + wm * (BM / 2) + a * 16 + rl;
    if (gi >= M) continue;
#pragma unroll
    for (int b = 0; b < WNT; b++) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int gj = n0 + wn * (BN / 2) + b * 16 + (F32 ? 4 * cg + v : cg + 4 * v);
        if (gj >= N || (c_upper && gi > gj)) continue;
        double r = (double)acc[a][b][v];
        if (P) {
          P[(size_t)blockIdx.z * M * N + (size_t)gi + (size_t)gj * M] = r;
        } else {
          double o = alpha * r;
          if (read_c && gi >= zero_rows) o += beta * cv[a][b][v];
          C[(size_t)gi + (size_t)gj * ldc] = o;
        }
      }
    }
  }
}

// 64 outputs per workgroup; wave q sums partials q, q + 4, ... with two
// accumulators (several loads in flight), the four wave sums are combined in
// a fixed order: deterministic for a given split count
__global__ __launch_bounds__(256) void splitk_reduce_kernel(int M, int N, int S, double alpha,
                                                            double beta,
                                                            const double *__restrict__ P,
                                                            double *__restrict__ C, int ldc,
                                                            int zero_rows, int c_upper) {
  __shared__ double red[4][64];
  const int l = threadIdx.x & 63, q = threadIdx.x >> 6;
  const size_t tot = (size_t)M * N;
  const size_t idx = (size_t)blockIdx.x * 64 + l;
  double s = 0.0;
  if (idx < tot) {
    const double *p = P + idx;
    double s0 = 0.0, s1 = 0.0;
    int z = q;
    for (; z + 12 < S; z += 16) {
      const double a0 = p[(size_t)z * tot], a1 = p[(size_t)(z + 4) * tot];
      const double a2 = p[(size_t)(z + 8) * tot], a3 = p[(size_t)(z + 12) * tot];
      s0 += a0 + a2;
      s1 += a1 + a3;
    }
    for (; z < S; z += 4) s0 += p[(size_t)z * tot];
    s = s0 + s1;
  }
  red[q][l] = s;
  __syncthreads();
  if (q != 0 || idx >= tot) return;
  const int i = (int)(idx % M), j = (int)(idx / M);
  if (c_upper && i > j) return;
  s = (red[0][l] + red[1][l]) + (red[2][l] + red[3][l]);
  double *c = &C[(size_t)i + (size_t)j * ldc];
  double o = alpha * s;
  if (beta != 0.0 && i >= zero_rows) o += beta * (*c);
  *c = o;
}

// W(:, j) = alpha op(T) sum_z P_z(:, j) for a split product of m <= MAXM
// rows (op(T) = T^T when transT, else T; T upper): one workgroup per column
// j.  Thread t sums partials z = q, q + Q, ... of row i (i = t % Mp,
// q = t / Mp, Mp = 32, 64 or 128, Q = 256 / Mp) with four accumulators; the Q sums are
// combined in a fixed order, so the result does not depend on timing.
// op(T) is staged in LDS.
template <int MAXM>
__global__ __launch_bounds__(256) void splitk_reduce_tmul_kernel(int M, int N, int S, double alpha,
                                                                 const double *__restrict__ P,
                                                                 const double *__restrict__ T,
                                                                 int ldt, int transT,
                                                                 double *__restrict__ W, int ldw) {
  __shared__ double red[256];
  __shared__ double g[MAXM];
  __shared__ double Ts[MAXM][MAXM + 1];
  const int j = blockIdx.x, t = threadIdx.x;
  const int Mp = (M <= 32) ? 32 : (M <= 64) ? 64 : 128, Q = 256 / Mp;
  const int i = t % Mp, q = t / Mp;
  const size_t tot = (size_t)M * N;
  for (int e = t; e < M * M; e += 256) {
    const int r = e % M, c = e / M;
    double v = 0.0;
    if (transT) {
      if (c <= r) v = T[(size_t)c + (size_t)r * ldt];
    } else if (c >= r) {
      v = T[(size_t)r + (size_t)c * ldt];
    }
    Ts[r][c] = v;
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  if (i < M) {
    const double *p = P + i + (size_t)j * M;
    int z = q;
    for (; z + 3 * Q < S; z += 4 * Q) {
      s0 += p[(size_t)z * tot];
      s1 += p[(size_t)(z + Q) * tot];
      s2 += p[(size_t)(z + 2 * Q) * tot];
      s3 += p[(size_t)(z + 3 * Q) * tot];
    }
    for (; z < S; z += Q) s0 += p[(size_t)z * tot];
  }
  red[q * Mp + i] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (t < M) {
    double s = red[t];
    for (int r = 1; r < Q; r++) s += red[r * Mp + t];
    g[t] = s;
  }
  __syncthreads();
  if (t < M) {
    double w = 0.0;
    for (int l = 0; l < M; l++) w += Ts[t][l] * g[l];
    W[(size_t)t + (size_t)j * ldw] = alpha * w;
  }
}

template <int BM, int BN>
void launch_tile(const ld_gemm_t *g, int splits, int kchunk, double *P, hipStream_t st) {
  dim3 grid((g->m + BM - 1) / BM, (g->n + BN - 1) / BN, splits);
  dim3 block(NT);
  char cls[96];
  const bool f32 = ld_tune()->gemm_lowp != 0;
  snprintf(cls, sizeof cls, "%s<%d, %d, %s, %s>", f32 ? "gemm_f32mfma_kernel" : "gemm_f64_kernel", BM,
           BN, g->ta ? "true" : "false", g->tb ? "true" : "false");
  const double M = g->m, N = g->n, K = g->k;
  const int ph = ld_prof_begin(cls);
#define LGO(TA_, TB_)                                                                      \
  do {                                                                                     \
    if (f32)                                                                               \
      hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, TA_, TB_, true>), grid, block, 0, st,    \
                         g->m, g->n, g->k, g->alpha, g->beta, g->A, g->lda, g->sa, g->B,   \
                         g->ldb, g->sb, g->C, g->ldc, g->zero_rows, g->c_upper, kchunk, P); \
    else                                                                                   \
      hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, TA_, TB_>), grid, block, 0, st, g->m,    \
                         g->n, g->k, g->alpha, g->beta, g->A, g->lda, g->sa, g->B, g->ldb,  \
                         g->sb, g->C, g->ldc, g->zero_rows, g->c_upper, kchunk, P);        \
  } while (0)
  if (!g->ta && !g->tb) LGO(false, false);
  else if (!g->ta && g->tb) LGO(false, true);
  else if (g->ta && !g->tb) LGO(true, false);
  else LGO(true, true);
#undef LGO
  ld_prof_end(ph, 2.0 * M * N * K,
              8.0 * (M * K + K * N + (P ? (double)splits * M * N : 2.0 * M * N)));
}

// ---------------------------------------------------------------------------
// 128 x 128 tiles with direct-to-LDS operand loads (global_load_lds_dwordx4:
// 16 bytes per lane go from HBM / L2 straight into LDS, no VGPR staging), for
// general operands of products op(A) B with op(A) = A or A^T, B not
// transposed, every dimension a multiple of the tile (128, 128, 16) and 16-byte
// aligned operands (the dispatcher below splits structured edges and ragged
// strips off to the kernel above).  Two LDS stages, one barrier per k step.
// LDS layouts, chosen so the MFMA operand reads are free of bank conflicts:
//   k-contiguous operands (B; A when transposed): column x's 16 k values as
//     8 16-byte chunks (k, k + 1) at x * 16 doubles, chunk c in slot
//     c ^ ((x >> 1) & 7); a load instruction fills 8 whole columns, 8 lanes
//     per column reading its 128 contiguous bytes;
//   m-contiguous A: [k][144], a 1 KiB column per load instruction, the
//     16-double pad putting k and k + 1 on opposite halves of the banks.
// 256 threads = 2 x 2 waves of 64 x 64 (4 x 4 MFMA 16x16x4 tiles each).
// ---------------------------------------------------------------------------
constexpr int G2_BM = 128, G2_BN = 128, G2_BK = 16;
constexpr int G2_ASTAGE = 16 * 144;  // doubles per A stage (the transposed layout uses 2048)
constexpr int G2_BSTAGE = 2048;
constexpr size_t G2_LDS = sizeof(double) * 2 * (G2_ASTAGE + G2_BSTAGE);

typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ void g2_load(const double *gp, double *lp) {
  __builtin_amdgcn_global_load_lds((const void *)gp, (lds_void *)lp, 16, 0, 0);
}

template <bool TA, bool KUP = false>
__global__ __launch_bounds__(NT, 2) void gemm2_f64_kernel(int M, int N, int K, double alpha, double beta,
                                                         const double *__restrict__ A, int lda,
                                                         const double *__restrict__ B, int ldb,
                                                         double *__restrict__ C, int ldc, int kchunk,
                                                         double *__restrict__ P, int kupper) {
  extern __shared__ __attribute__((aligned(16))) double g2s[];
  double *As = g2s;                   // [2][G2_ASTAGE]
  double *Bs = g2s + 2 * G2_ASTAGE;   // [2][G2_BSTAGE]
  int tx = blockIdx.x, ty = blockIdx.y;
  {
    // XCD-aware tile order (as in the kernel above)
    const int gx = gridDim.x, gy = gridDim.y, ntile = gx * gy;
    if (ntile >= 512 && gx % 8 == 0 && gy % 8 == 0) {
      const int lin = blockIdx.x + blockIdx.y * gx;
      const int t = (lin % 8) * (ntile / 8) + lin / 8;
      const int st = t / 64, w = t % 64, sx = gx / 8;
      tx = (st % sx) * 8 + w % 8;
      ty = (st / sx) * 8 + w / 8;
    }
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = tx * G2_BM, n0 = ty * G2_BN;
  const int kbeg = blockIdx.z * kchunk;
  // first column (within the tile) of the wave's 16-column tile b: 64
  // contiguous columns per wave, or with KUP the two waves' tiles
  // interleaved (0, 32, 64, 96 and 16, 48, 80, 112), so that both meet the
  // triangle's skipped tiles alike and neither waits for the other
  auto g2_col = [&](int wnn, int b) { return KUP ? wnn * 16 + b * 32 : wnn * 64 + b * 16; };
  // an upper triangular B with zeros below its diagonal: rows of B past the
  // tile's last column add nothing
  const int kend = min(kupper ? min(K, n0 + G2_BN) : K, kbeg + kchunk);
  const int nsteps = max(0, (kend - kbeg) / G2_BK);

  // this wave's four load instructions per operand and stage; load jj of a
  // k-contiguous operand fills columns 8 jj .. 8 jj + 7
  auto issue = [&](int buf, int k0) {
    double *as = As + buf * G2_ASTAGE, *bs = Bs + buf * G2_BSTAGE;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int jj = wave * 4 + j;
      const int x = jj * 8 + (lane >> 3), c = (lane & 7) ^ ((x >> 1) & 7);
      g2_load(B + (size_t)(k0 + 2 * c) + (size_t)(n0 + x) * ldb, bs + jj * 128);
      if (TA)
        g2_load(A + (size_t)(k0 + 2 * c) + (size_t)(m0 + x) * lda, as + jj * 128);
      else
        g2_load(A + (size_t)(m0 + 2 * lane) + (size_t)(k0 + jj) * lda, as + jj * 144);
    }
  };

  d4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = d4_t{0.0, 0.0, 0.0, 0.0};

  const int r = lane & 15, kk = lane >> 4;
  if (nsteps > 0) issue(0, kbeg);
  for (int t = 0; t < nsteps; t++) {
    __builtin_amdgcn_s_waitcnt(0);  // this wave's loads of step t are in LDS
    __syncthreads();                // everyone's, and step t - 1's stage is free
    if (t + 1 < nsteps) issue((t + 1) & 1, kbeg + (t + 1) * G2_BK);
    const double *as = As + (t & 1) * G2_ASTAGE, *bs = Bs + (t & 1) * G2_BSTAGE;
    // operand fragments of sub-step s + 1 read from LDS while the 16 MFMAs
    // of sub-step s issue
    double af[2][4], bf[2][4];
    auto frags = [&](int sstep, double (&fa)[4], double (&fb)[4]) {
      const int k = 4 * sstep + kk;
#pragma unroll
      for (int a = 0; a < 4; a++) {
        const int m = wm * 64 + a * 16 + r;
        fa[a] = TA ? as[m * 16 + (((k >> 1) ^ ((m >> 1) & 7)) << 1) + (k & 1)] : as[k * 144 + m];
      }
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int n = g2_col(wn, b) + r;
        fb[b] = bs[n * 16 + (((k >> 1) ^ ((n >> 1) & 7)) << 1) + (k & 1)];
      }
    };
    frags(0, af[0], bf[0]);
#pragma unroll
    for (int sstep = 0; sstep < G2_BK / 4; sstep++) {
      const int cur = sstep & 1;
      if (sstep + 1 < G2_BK / 4) frags(sstep + 1, af[cur ^ 1], bf[cur ^ 1]);
#pragma unroll
      for (int b = 0; b < 4; b++) {
        // KUP: B upper triangular with zeros below the diagonal: rows of B
        // past a 16-column tile's last column add nothing to it, so the wave
        // skips that tile's MFMAs from there (a wave-uniform test; the freed
        // MFMA cycles go to the CU's other workgroup)
        if (KUP && kbeg + t * G2_BK + 4 * sstep > n0 + g2_col(wn, b) + 15) continue;
#pragma unroll
        for (int a = 0; a < 4; a++) acc[a][b] = mfma16(bf[cur][b], af[cur][a], acc[a][b]);
      }
    }
  }

  // epilogue (every tile is interior): lane holds C(row m0 + .. + r,
  // column n0 + .. + kk + 4v); by 16-row groups, each group's 16 C loads
  // issued before its first store (the registers of all 64 would spill)
  const bool read_c = (P == nullptr) && beta != 0.0;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int gi = m0 + wm * 64 + a * 16 + r;
    double cv[4][4];
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int gj = n0 + g2_col(wn, b) + kk + 4 * v;
        cv[b][v] = read_c ? C[(size_t)gi + (size_t)gj * ldc] : 0.0;
      }
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int gj = n0 + g2_col(wn, b) + kk + 4 * v;
        if (P) {
          P[(size_t)blockIdx.z * M * N + (size_t)gi + (size_t)gj * M] = acc[a][b][v];
        } else {
          double o = alpha * acc[a][b][v];
          if (read_c) o += beta * cv[b][v];
          C[(size_t)gi + (size_t)gj * ldc] = o;
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Small products (M, N <= 512, K <= 512): the products of the recursion's
// block steps with a T factor (T1^T W, T1 X, H T2 at 128..512 columns) are
// a few to a few hundred megaflops each, and the tiled kernel's k loop makes
// them a chain of global-load latencies (12-17 us, plus a split-K reduction
// at 512).  Here a workgroup owns a 32 x 32 tile of C and stages op(A)'s 32
// rows and op(B)'s 32 columns for up to 256 k at once (one batch of loads
// per operand, the structure masks applied on the way), then each wave runs
// the MFMA chain of its 16 x 16 quarter; K > 256 takes a second round.
// LDS: [k][32] per operand (row stride SG_LD), read conflict-free by the
// MFMA operand layout (16 consecutive doubles per k).
// ---------------------------------------------------------------------------
constexpr int SG_T = 32;          // output tile
constexpr int SG_KC = 256;        // k per LDS round
constexpr int SG_KMAX = 512;      // K of the products taken
constexpr int SG_MAX = 512;       // M, N of the products taken
constexpr int SG_LD = SG_T + 2;   // LDS row stride (doubles)
constexpr int SG_LDS = 2 * SG_KC * SG_LD * 8;

template <bool TA, bool TB>
__global__ __launch_bounds__(NT) void gemm_small_kernel(int M, int N, int K, double alpha, double beta,
                                                        const double *__restrict__ A, int lda, int sa,
                                                        const double *__restrict__ B, int ldb, int sb,
                                                        double *__restrict__ C, int ldc, int zero_rows) {
  extern __shared__ double sg_lds[];
  double *As = sg_lds, *Bs = sg_lds + SG_KC * SG_LD;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, li = lane & 15, lk = lane >> 4;
  const int wi = wave & 1, wj = wave >> 1;
  const int i0 = blockIdx.x * SG_T, j0 = blockIdx.y * SG_T;
  d4_t acc = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += SG_KC) {
    const int kc = min(SG_KC, K - k0);
    if (k0 > 0) __syncthreads();  // the previous round's reads are done
    double va[SG_T * SG_KC / NT], vb[SG_T * SG_KC / NT];
#pragma unroll
    for (int u = 0; u < SG_T * SG_KC / NT; u++) {
      const int e = t + NT * u;
      // op(A)(i, k): along i when A is stored M x K, along k when K x M
      const int ia = TA ? (e >> 8) : (e & 31), ka = TA ? (e & 255) : (e >> 5);
      const int gi = i0 + ia, gk = k0 + ka;
      va[u] = (gi < M && ka < kc) ? (TA ? mask_val(A, lda, gk, gi, sa) : mask_val(A, lda, gi, gk, sa)) : 0.0;
      // op(B)(k, j): along k when B is stored K x N, along j when N x K
      const int jb = TB ? (e & 31) : (e >> 8), kb = TB ? (e >> 5) : (e & 255);
      const int gj = j0 + jb, gkb = k0 + kb;
      vb[u] = (gj < N && kb < kc) ? (TB ? mask_val(B, ldb, gj, gkb, sb) : mask_val(B, ldb, gkb, gj, sb)) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < SG_T * SG_KC / NT; u++) {
      const int e = t + NT * u;
      const int ia = TA ? (e >> 8) : (e & 31), ka = TA ? (e & 255) : (e >> 5);
      const int jb = TB ? (e & 31) : (e >> 8), kb = TB ? (e >> 5) : (e & 255);
      As[ka * SG_LD + ia] = va[u];
      Bs[kb * SG_LD + jb] = vb[u];
    }
    __syncthreads();
    const int steps = (kc + 3) >> 2;
    for (int s = 0; s < steps; s++) {
      const int k = 4 * s + lk;
      acc = mfma16(As[k * SG_LD + 16 * wi + li], Bs[k * SG_LD + 16 * wj + li], acc);
    }
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int r = i0 + 16 * wi + lk + 4 * v, c = j0 + 16 * wj + li;
    if (r < M && c < N) {
      double *cp = &C[(size_t)r + (size_t)c * ldc];
      double o = alpha * acc[v];
      if (beta != 0.0 && r >= zero_rows) o += beta * (*cp);
      *cp = o;
    }
  }
}

}  // namespace

// tile shape and split-K count of a product
static void gemm_plan(const ld_gemm_t *g, int &bm, int &bn, int &splits, int &kchunk) {
  const int M = g->m, N = g->n, K = g->k;
  const ld_tune_t *tu = ld_tune();
  bm = g->ta ? tu->gemm_bm_t : tu->gemm_bm;
  bn = g->ta ? tu->gemm_bn_t : tu->gemm_bn;
  if (M <= 32) bm = 32; else if (M <= 64 && bm > 64) bm = 64;
  if (N <= 32) bn = 32; else if (N <= 64 && bn > 64) bn = 64;
  long tiles = (long)((M + bm - 1) / bm) * ((N + bn - 1) / bn);
  // too few tiles to fill the chip: step down to 64 x 128 (from a larger
  // tile), then to 64 x 64
  const int small = g->ta ? tu->gemm_small_tiles_t : tu->gemm_small_tiles;
  if (tiles < small && bm * bn > 64 * 128 && N > 64) {
    bm = 64; bn = 128;
    tiles = (long)((M + bm - 1) / bm) * ((N + bn - 1) / bn);
  }
  if (tiles < small && bm * bn > 64 * 64) {
    bm = 64; bn = 64;
    tiles = (long)((M + bm - 1) / bm) * ((N + bn - 1) / bn);
  }
  splits = 1;
  const int target = tu->gemm_min_wgs;
  if (tiles < target && K >= 256) {
    splits = (int)((target + tiles - 1) / tiles);
    // down to kmin rows of K per split (4 k-steps of BK by default): a
    // long-K product with a one-tile output then spreads over up to K/kmin
    // workgroups
    const int kmin = (tu->gemm_kmin >= BK) ? tu->gemm_kmin : 64;
    int maxs = K / kmin;
    if (splits > maxs) splits = maxs;
    // cap partial storage at 64 Mi doubles
    long cap = (64L << 20) / ((long)M * N);
    if (splits > cap) splits = (int)cap;
    if (splits < 1) splits = 1;
  }
  kchunk = K;
  if (splits > 1) {
    kchunk = ((K + splits - 1) / splits + BK - 1) / BK * BK;
    splits = (K + kchunk - 1) / kchunk;
  }
}

static void dispatch_tile(const ld_gemm_t *g, int bm, int bn, int splits, int kchunk, double *P,
                          hipStream_t st) {
#define DISPATCH(BM_, BN_) if (bm == BM_ && bn == BN_) launch_tile<BM_, BN_>(g, splits, kchunk, P, st)
  DISPATCH(128, 128); else DISPATCH(128, 64); else DISPATCH(128, 32);
  else DISPATCH(64, 128); else DISPATCH(64, 64); else DISPATCH(64, 32);
  else DISPATCH(32, 128); else DISPATCH(32, 64); else DISPATCH(32, 32);
#undef DISPATCH
}

static void gemm_tiles(const ld_gemm_t *g);

// the 128 x 128 direct-to-LDS kernel on a general product (sa = sb = GEN,
// no transposed B, 16-byte aligned operands): its 128-aligned interior with
// K a multiple of 16, the ragged strips and the K tail by the tiled kernel
static bool g2_eligible(const ld_gemm_t *g) {
  // (products with op(A) = A^T run on it only with gemm_v2 = 2: in the
  // factorization they measured no faster than on the tiled kernel)
  return (!g->ta || ld_tune()->gemm_v2 >= 2) && g->sa == LD_GEN &&
         (g->sb == LD_GEN || g->sb == LD_UPPER_Z) && !g->tb &&
         !g->c_upper && g->zero_rows == 0 && !ld_tune()->gemm_lowp && ld_tune()->gemm_v2 &&
         ((uintptr_t)g->A & 15) == 0 && ((uintptr_t)g->B & 15) == 0 && (g->lda & 1) == 0 &&
         (g->ldb & 1) == 0;
}

// inner-product length summed over the columns of a product with an
// upper triangular B whose K range is capped per 128-column tile
static double flops_upper_cols(int Ni, int Ki) {
  double s = 0.0;
  for (int n0 = 0; n0 < Ni; n0 += G2_BN) s += (double)G2_BN * (n0 + G2_BN < Ki ? n0 + G2_BN : Ki);
  return s;
}

// A plain NN product (no structure, no transposes, no zero rows): the
// interior the dispatcher sends to the direct-to-LDS kernel goes to rocBLAS
// instead when the backend setting asks for it (lila_set_gemm_backend(1);
// rocBLAS's dgemm measured 69-70 TFLOP/s on the trailing-update shapes
// where gemm2_f64_kernel reaches 63-66).  False when not taken.
static bool gemm_rocblas(const ld_gemm_t *g) {
  static int state = 0;  // 0 not tried, 1 handle ready, -1 unavailable
  static rocblas_handle hd = nullptr;
  if (!ld_tune()->gemm_rocblas || g->sa != LD_GEN || g->sb != LD_GEN || g->zero_rows || g->c_upper || g->ta ||
      g->tb)
    return false;
  if (state == 0) {
    state = (rocblas_create_handle(&hd) == rocblas_status_success) ? 1 : -1;
    if (state < 0) fprintf(stderr, "lila_hip: no rocBLAS handle; plain products stay on gemm2_f64_kernel\n");
  }
  if (state < 0 || rocblas_set_stream(hd, (hipStream_t)ld_stream()) != rocblas_status_success) return false;
  const int ph = ld_prof_begin("rocblas_dgemm");
  const rocblas_status st = rocblas_dgemm(hd, rocblas_operation_none, rocblas_operation_none, g->m, g->n, g->k,
                                          &g->alpha, g->A, g->lda, g->B, g->ldb, &g->beta, g->C, g->ldc);
  ld_prof_end(ph, 2.0 * g->m * g->n * (double)g->k,
              8.0 * ((double)g->m * g->k + (double)g->k * g->n + 2.0 * g->m * g->n));
  if (st != rocblas_status_success) {
    fprintf(stderr, "lila_hip: rocblas_dgemm failed (%d)\n", (int)st);
    ld_set_info(-1);
  }
  return true;
}

static void gemm2_general(const ld_gemm_t *g) {
  if (gemm_rocblas(g)) return;
  const int M = g->m, N = g->n, K = g->k;
  const int Mi = M / G2_BM * G2_BM, Ni = N / G2_BN * G2_BN, Ki = K / G2_BK * G2_BK;
  const long tiles = (long)(Mi / G2_BM) * (Ni / G2_BN);
  if (tiles < 128 || Ki == 0) {
    gemm_tiles(g);
    return;
  }
  hipStream_t st = (hipStream_t)ld_stream();
  static int attr = 0;
  if (!attr) {
    for (const void *f : {(const void *)gemm2_f64_kernel<false>, (const void *)gemm2_f64_kernel<true>,
                          (const void *)gemm2_f64_kernel<false, true>}) {
      if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)G2_LDS) != hipSuccess) {
        (void)hipGetLastError();
        ld_tune()->gemm_v2 = 0;
        gemm_tiles(g);
        return;
      }
    }
    attr = 1;
  }
  {
    // split K when the tiles give fewer workgroups than the split-K target
    // (512 by default: two per CU, the kernel's occupancy), down to 256 k
    // per split: the partials are summed and scaled by the split-K reduction.
    // From three quarters of the target on, the tiles alone nearly fill the
    // chip and the reduction's pass over the partials (2 M N doubles and
    // more) costs more than the split gains
    int splits = 1, kchunk = Ki;
    const long target = ld_tune()->gemm_min_wgs;
    if (tiles < target && 4 * tiles < 3 * target && Ki >= 512) {
      splits = (int)((target + tiles - 1) / tiles);
      if (splits > Ki / 256) splits = Ki / 256;
      const long cap = (64L << 20) / ((long)Mi * Ni);
      if (splits > cap) splits = (int)cap;
      if (splits < 1) splits = 1;
      kchunk = ((Ki + splits - 1) / splits + G2_BK - 1) / G2_BK * G2_BK;
      splits = (Ki + kchunk - 1) / kchunk;
    }
    double *P = nullptr;
    size_t mark = 0;
    if (splits > 1) {
      mark = ld_ws_mark();
      P = (double *)ld_ws_alloc((size_t)splits * Mi * Ni * sizeof(double));
      if (ld_failed()) {
        ld_ws_release(mark);
        return;
      }
    }
    const int ph = ld_prof_begin(g->ta ? "gemm2_f64_kernel<true>" : "gemm2_f64_kernel<false>");
    dim3 grid(Mi / G2_BM, Ni / G2_BN, splits);
    const int kup = (g->sb == LD_UPPER_Z) ? 1 : 0;
    if (g->ta)
      hipLaunchKernelGGL(gemm2_f64_kernel<true>, grid, dim3(NT), G2_LDS, st, Mi, Ni, Ki, g->alpha, g->beta,
                         g->A, g->lda, g->B, g->ldb, g->C, g->ldc, kchunk, P, kup);
    else if (kup)
      hipLaunchKernelGGL((gemm2_f64_kernel<false, true>), grid, dim3(NT), G2_LDS, st, Mi, Ni, Ki, g->alpha,
                         g->beta, g->A, g->lda, g->B, g->ldb, g->C, g->ldc, kchunk, P, kup);
    else
      hipLaunchKernelGGL(gemm2_f64_kernel<false>, grid, dim3(NT), G2_LDS, st, Mi, Ni, Ki, g->alpha, g->beta,
                         g->A, g->lda, g->B, g->ldb, g->C, g->ldc, kchunk, P, kup);
    ld_prof_end(ph, kup ? 2.0 * Mi * flops_upper_cols(Ni, Ki) : 2.0 * Mi * Ni * Ki,
                8.0 * ((double)Mi * Ki + (double)Ki * Ni + (splits > 1 ? (double)splits : 2.0) * Mi * Ni));
    if (splits > 1) {
      const size_t tot = (size_t)Mi * Ni;
      const int ph2 = ld_prof_begin("splitk_reduce_kernel");
      hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)((tot + 63) / 64)), dim3(256), 0, st, Mi, Ni,
                         splits, g->alpha, g->beta, (const double *)P, g->C, g->ldc, 0, 0);
      ld_prof_end(ph2, (double)splits * tot, 8.0 * ((double)splits * tot + 2.0 * tot));
      ld_ws_release(mark);
    }
  }
  // the K tail onto the interior (beta = 1 for every row); with LD_UPPER_Z
  // the strips read B's stored zeros as a general operand
  if (Ki < K) {
    ld_gemm_t h = *g;
    if (h.sb == LD_UPPER_Z) h.sb = LD_GEN;
    h.m = Mi; h.n = Ni; h.k = K - Ki; h.beta = 1.0; h.zero_rows = 0;
    h.A = g->ta ? g->A + Ki : g->A + (size_t)Ki * g->lda;
    h.B = g->B + Ki;
    gemm_tiles(&h);
  }
  // right strip (all rows), bottom strip
  if (Ni < N) {
    ld_gemm_t h = *g;
    if (h.sb == LD_UPPER_Z) h.sb = LD_GEN;
    h.n = N - Ni;
    h.B = g->B + (size_t)Ni * g->ldb;
    h.C = g->C + (size_t)Ni * g->ldc;
    gemm_tiles(&h);
  }
  if (Mi < M) {
    ld_gemm_t h = *g;
    if (h.sb == LD_UPPER_Z) h.sb = LD_UPPER;
    h.m = M - Mi; h.n = Ni;
    h.A = g->ta ? g->A + (size_t)Mi * g->lda : g->A + Mi;
    h.C = g->C + Mi;
    gemm_tiles(&h);
  }
}

extern "C" void ld_gemm(const ld_gemm_t *g) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (g->m <= 0 || g->n <= 0) return;
  if (g->k > 0 && g->alpha != 0.0 && ld_tune()->gemm_v2 && !ld_tune()->gemm_lowp && !g->tb &&
      (g->sb == LD_GEN || (g->sb == LD_UPPER_Z && g->sa == LD_GEN)) && !g->c_upper &&
      (g->sa == LD_GEN || g->sa == LD_VUNIT)) {
    // the general part of the product (all of it, or what is left of a
    // Householder V once its unit lower top block is split off: C's first k
    // rows when op(A) = V, the first m rows of K when op(A) = V^T)
    ld_gemm_t h = *g, top = *g;
    bool split = false;
    if (g->sa == LD_VUNIT) {
      split = true;
      if (!g->ta) {
        const int kv = g->k;
        if (g->zero_rows > kv || g->m <= kv) split = false;
        top.m = kv;
        h.m = g->m - kv; h.A = g->A + kv; h.C = g->C + kv; h.sa = LD_GEN; h.zero_rows = 0;
      } else {
        const int kv = g->m;
        if (g->k <= kv || g->zero_rows > 0) split = false;
        top.k = kv;
        h.k = g->k - kv; h.A = g->A + kv; h.B = g->B + kv; h.sa = LD_GEN; h.beta = 1.0;
      }
      if (!split) {
        gemm_tiles(g);
        return;
      }
    }
    const long tiles = (long)(h.m / G2_BM) * (h.n / G2_BN);
    if (g2_eligible(&h) && tiles >= 128 && h.k >= G2_BK) {
      if (split) gemm_tiles(&top);
      gemm2_general(&h);
      return;
    }
  }
  gemm_tiles(g);
}

// the small-product kernel for M, N <= 512, K <= 512 (FP64, no SYRK form);
// false when the product is not one it takes (LILA_GEMM_SMALL=0 turns it off)
static bool gemm_small(const ld_gemm_t *g, hipStream_t st) {
  static int on = -1;
  if (on < 0) {
    const char *e = getenv("LILA_GEMM_SMALL");
    on = e ? (atoi(e) != 0) : 1;
    if (on) {
      for (const void *f : {(const void *)gemm_small_kernel<false, false>, (const void *)gemm_small_kernel<false, true>,
                            (const void *)gemm_small_kernel<true, false>, (const void *)gemm_small_kernel<true, true>})
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, SG_LDS) != hipSuccess) {
          (void)hipGetLastError();
          on = 0;
        }
    }
  }
  if (!on || g->m > SG_MAX || g->n > SG_MAX || g->k > SG_KMAX || g->c_upper || ld_tune()->gemm_lowp) return false;
  const dim3 grid((g->m + SG_T - 1) / SG_T, (g->n + SG_T - 1) / SG_T);
  const int ph = ld_prof_begin("gemm_small_kernel");
#define SGO(TA_, TB_)                                                                                     \
  hipLaunchKernelGGL((gemm_small_kernel<TA_, TB_>), grid, dim3(NT), SG_LDS, st, g->m, g->n, g->k, g->alpha, \
                     g->beta, g->A, g->lda, g->sa, g->B, g->ldb, g->sb, g->C, g->ldc, g->zero_rows)
  if (!g->ta && !g->tb) SGO(false, false);
  else if (!g->ta && g->tb) SGO(false, true);
  else if (g->ta && !g->tb) SGO(true, false);
  else SGO(true, true);
#undef SGO
  const double M = g->m, N = g->n, K = g->k;
  ld_prof_end(ph, 2.0 * M * N * K, 8.0 * (M * K + K * N + 2.0 * M * N));
  return true;
}

static void gemm_tiles(const ld_gemm_t *g0) {
  if (g0->m <= 0 || g0->n <= 0) return;
  // the tiled kernel knows LD_UPPER_Z as LD_UPPER (it masks the lower part)
  ld_gemm_t gz = *g0;
  if (gz.sa == LD_UPPER_Z) gz.sa = LD_UPPER;
  if (gz.sb == LD_UPPER_Z) gz.sb = LD_UPPER;
  const ld_gemm_t *g = &gz;
  hipStream_t st = (hipStream_t)ld_stream();
  if (g->k <= 0 || g->alpha == 0.0) {
    // C = beta C (rows < zero_rows -> 0)
    ld_gemm_t h = *g;
    h.k = 0;
    launch_tile<64, 64>(&h, 1, 0, nullptr, st);
    return;
  }
  const int M = g->m, N = g->n;
  if (gemm_small(g, st)) return;
  int bm, bn, splits, kchunk;
  gemm_plan(g, bm, bn, splits, kchunk);
  double *P = nullptr;
  size_t mark = 0;
  if (splits > 1) {
    mark = ld_ws_mark();
    P = (double *)ld_ws_alloc((size_t)splits * M * N * sizeof(double));
  }
  dispatch_tile(g, bm, bn, splits, kchunk, P, st);
  if (splits > 1) {
    size_t tot = (size_t)M * N;
    const int ph = ld_prof_begin("splitk_reduce_kernel");
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)((tot + 63) / 64)), dim3(256), 0, st,
                       M, N, splits, g->alpha, g->beta, P, g->C, g->ldc, g->zero_rows,
                       g->c_upper);
    ld_prof_end(ph, (double)splits * tot, 8.0 * ((double)splits * tot + 2.0 * tot));
    ld_ws_release(mark);
  }
}

// W (m x n, ldw) = op(T) (alpha op(A) op(B)) with op(T) = T^T (transT) or T,
// T upper m x m: the split-K partials of the product go straight into one
// launch that sums each column and multiplies it by op(T).  Returns 1 and
// does nothing when the product would not be split, m > 64 or the GEMMs run
// in FP32 (the caller then runs the product and the triangular multiply as
// two GEMMs).  beta, C, zero_rows and c_upper of g are not used.
extern "C" int ld_gemm_tmul(const ld_gemm_t *g, const double *T, int ldt, int transT, double *W,
                            int ldw) {
  if (ld_failed()) return 0;  // an earlier allocation of this call failed
  // (m up to 128 measured no faster: two summing groups per row, 130 KB of LDS)
  if (g->m <= 0 || g->n <= 0 || g->k <= 0 || g->alpha == 0.0 || g->m > 64 || ld_tune()->gemm_lowp)
    return 1;
  int bm, bn, splits, kchunk;
  gemm_plan(g, bm, bn, splits, kchunk);
  if (splits <= 1) return 1;
  hipStream_t st = (hipStream_t)ld_stream();
  const int M = g->m, N = g->n;
  const size_t mark = ld_ws_mark();
  double *P = (double *)ld_ws_alloc((size_t)splits * M * N * sizeof(double));
  dispatch_tile(g, bm, bn, splits, kchunk, P, st);
  const int ph = ld_prof_begin("splitk_reduce_tmul_kernel");
  hipLaunchKernelGGL(splitk_reduce_tmul_kernel<64>, dim3(N), dim3(256), 0, st, M, N, splits,
                     g->alpha, P, T, ldt, transT, W, ldw);
  ld_prof_end(ph, (double)splits * M * N + (double)M * M * N,
              8.0 * ((double)splits * M * N + (double)M * N));
  ld_ws_release(mark);
  return 0;
}
