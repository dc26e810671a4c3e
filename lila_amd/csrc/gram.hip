// gram.hip -- the Gram matrix G = A^T A (upper triangle, n <= 256) of a
// tall FP64 matrix, in FP64 MFMA arithmetic (gram64_kernel, at the end) or
// (gram3_kernel) with 3-term split-BF16 MFMA products and FP32
// accumulation: each entry a = hi + mid + lo with three BF16 terms (the
// residuals a - hi and a - hi - mid are exact in FP64), and
// a^T b ~= hi^T hi' + hi^T mid' + mid^T hi' + hi^T lo' + lo^T hi' + mid^T mid'
// (the dropped terms are ~2^-24 |a||b|).  It is the first-pass Gram of the
// tall-skinny Cholesky-QR2 path (one read of A: HBM-bound at ~8 B per entry).
//
// Layout: one workgroup of 512 threads (8 waves) per row chunk; a stage of
// 32 rows x n columns is loaded into registers two stages ahead, split, and
// written to one of two LDS buffers as Ts[buf][term][column][row] BF16 while
// the MFMAs read the other (one barrier per stage; row stride 40 halves: 16
// consecutive columns hit distinct banks); the 16 x 16 output tiles I <= J
// are dealt to the waves by row pairs (I = w and 15 - w) so each wave owns
// 17 tiles of a 16 x 16 tile grid and reuses its row operands across J.
// MFMA v_mfma_f32_16x16x32_bf16 (one per stage, term pair and tile): lane l
// holds A(row l%16, k 8(l/16)..+7), B(k 8(l/16)..+7, col l%16) and
// D(row 4(l/16) + v, col l%16).  Partials
// (FP64, one n x n per workgroup) are summed in a fixed order by a second
// kernel, so the result does not depend on timing.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lila_dev.h"

namespace {

constexpr int GNT = 512;   // threads per workgroup
constexpr int GNB = 256;   // largest n
constexpr int GBK = 32;    // rows per stage
constexpr int GLD = 40;    // LDS row stride of one column (halves)

typedef short bf4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef float f4v_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_rn(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ double bf16_val(unsigned short h) {
  return (double)__uint_as_float((unsigned)h << 16);
}

__global__ __launch_bounds__(GNT, 1) void gram3_kernel(int m, int n, const double *__restrict__ A,
                                                       int lda, int chunk,
                                                       double *__restrict__ P) {
  __shared__ unsigned short Ts[2][3][GNB][GLD];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r0 = blockIdx.x * chunk;
  const int r1 = min(m, r0 + chunk);
  const int nt = (n + 15) / 16;
  // stage loads: thread t covers column t / 2, rows 16 (t % 2) .. + 15
  const int lc = tid >> 1, lh = (tid & 1) * 16;
  // two stages in flight: st0 and st1 take alternate stages, each loaded
  // two stages before it is split into LDS
  double st0[16], st1[16];
  auto load_stage = [&](double (&st)[16], int rs) {
    // (a stage inside the chunk loads without per-element tests, so its 16
    // loads are in flight together)
    if (lc < n && rs + GBK <= r1) {
      const double *p = A + (size_t)(rs + lh) + (size_t)lc * lda;
#pragma unroll
      for (int q = 0; q < 16; q++) st[q] = p[q];
    } else {
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int r = rs + lh + q;
        st[q] = (lc < n && r < r1) ? A[(size_t)r + (size_t)lc * lda] : 0.0;
      }
    }
  };
  auto store_stage = [&](int b, const double (&st)[16]) {
#pragma unroll
    for (int q4 = 0; q4 < 16; q4 += 4) {
      bf4_t t0, t1, t2;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const double a = st[q4 + e];
        const unsigned short h = bf16_rn((float)a);
        double res = a - bf16_val(h);
        const unsigned short md = bf16_rn((float)res);
        res -= bf16_val(md);
        t0[e] = (short)h;
        t1[e] = (short)md;
        t2[e] = (short)bf16_rn((float)res);
      }
      *reinterpret_cast<bf4_t *>(&Ts[b][0][lc][lh + q4]) = t0;
      *reinterpret_cast<bf4_t *>(&Ts[b][1][lc][lh + q4]) = t1;
      *reinterpret_cast<bf4_t *>(&Ts[b][2][lc][lh + q4]) = t2;
    }
  };
  f4v_t acc[2][16];
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int j = 0; j < 16; j++) acc[s][j] = f4v_t{0.f, 0.f, 0.f, 0.f};
  const int rowt[2] = {w, 15 - w};
  const int col = lane & 15, kq = (lane >> 4) * 4;  // kq: output rows of this lane
  // one k = 32 MFMA per term pair and tile: lane l takes rows
  // 8 (l/16) .. + 7 of its column (16 bytes of LDS)
  const int k8 = (lane >> 4) * 8;
  auto compute_stage = [&](int b) {
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const int I = rowt[s];
      if (I >= nt) continue;
      bf8_t a3[3];
#pragma unroll
      for (int t = 0; t < 3; t++)
        a3[t] = *reinterpret_cast<const bf8_t *>(&Ts[b][t][I * 16 + col][k8]);
#pragma unroll
      for (int J = 0; J < 16; J++) {
        if (J < I || J >= nt) continue;
        bf8_t b3[3];
#pragma unroll
        for (int t = 0; t < 3; t++)
          b3[t] = *reinterpret_cast<const bf8_t *>(&Ts[b][t][J * 16 + col][k8]);
        f4v_t c = acc[s][J];
        // small terms first
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3[1], b3[1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3[0], b3[2], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3[2], b3[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3[0], b3[1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3[1], b3[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3[0], b3[0], c, 0, 0, 0);
        acc[s][J] = c;
      }
    }
  };

  // stage k lives in LDS buffer k % 2 and, before that, in register set
  // k % 2 (st0 / st1), loaded two stages ahead; each step splits stage k + 1
  // into LDS while the MFMAs read stage k, then one barrier
  if (r0 < r1) load_stage(st0, r0);
  if (r0 + GBK < r1) load_stage(st1, r0 + GBK);
  if (r0 < r1) store_stage(0, st0);
  __syncthreads();
  if (r0 + 2 * GBK < r1) load_stage(st0, r0 + 2 * GBK);
  for (int rs = r0; rs < r1; rs += 2 * GBK) {
    // stage rs in buffer 0, stage rs + GBK in st1
    if (rs + GBK < r1) store_stage(1, st1);
    if (rs + 3 * GBK < r1) load_stage(st1, rs + 3 * GBK);
    compute_stage(0);
    __syncthreads();
    if (rs + GBK >= r1) break;
    // stage rs + GBK in buffer 1, stage rs + 2 GBK in st0
    if (rs + 2 * GBK < r1) store_stage(0, st0);
    if (rs + 4 * GBK < r1) load_stage(st0, rs + 4 * GBK);
    compute_stage(1);
    __syncthreads();
  }
  // partial n x n (column-major) of this workgroup: tiles I <= J
  double *Pw = P + (size_t)blockIdx.x * n * n;
#pragma unroll
  for (int s = 0; s < 2; s++) {
    const int I = rowt[s];
    if (I >= nt) continue;
#pragma unroll
    for (int J = 0; J < 16; J++) {
      if (J < I || J >= nt) continue;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = I * 16 + kq + v, c = J * 16 + col;
        if (r < n && c < n) Pw[(size_t)r + (size_t)c * n] = (double)acc[s][J][v];
      }
    }
  }
}

// The same Gram in FP64: stages of 32 rows held in LDS as FP64 (column
// stride 34 doubles: the two k rows a half-wave reads fall in disjoint
// banks), v_mfma_f64_16x16x4_f64 (lane l: A(row l%16, k l/16),
// B(k l/16, col l%16), D(row l/16 + 4v, col l%16)), eight per tile per stage.
constexpr int GLD64 = 34;

typedef double d4v_t __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(GNT, 1) void gram64_kernel(int m, int n, const double *__restrict__ A,
                                                        int lda, int chunk,
                                                        double *__restrict__ P) {
  __shared__ double Ds[2][GNB][GLD64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r0 = blockIdx.x * chunk;
  const int r1 = min(m, r0 + chunk);
  const int nt = (n + 15) / 16;
  const int lc = tid >> 1, lh = (tid & 1) * 16;
  // one register stage: a stage's MFMA work (~7 us) covers the next load
  double st0[16];
  auto load_stage = [&](double (&st)[16], int rs) {
    if (lc < n && rs + GBK <= r1) {
      const double *p = A + (size_t)(rs + lh) + (size_t)lc * lda;
#pragma unroll
      for (int q = 0; q < 16; q++) st[q] = p[q];
    } else {
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int r = rs + lh + q;
        st[q] = (lc < n && r < r1) ? A[(size_t)r + (size_t)lc * lda] : 0.0;
      }
    }
  };
  auto store_stage = [&](int b, const double (&st)[16]) {
#pragma unroll
    for (int q = 0; q < 16; q += 2)
      *reinterpret_cast<double2 *>(&Ds[b][lc][lh + q]) = make_double2(st[q], st[q + 1]);
  };
  // the wave's 17 tiles, compactly: slot t < 16 - w is (w, w + t), the
  // rest (15 - w, 15 - w + t - (16 - w)); slots past n are idle
  d4v_t acc[17];
#pragma unroll
  for (int t = 0; t < 17; t++) acc[t] = d4v_t{0.0, 0.0, 0.0, 0.0};
  const int split = 16 - w, I0 = w, I1 = 15 - w;
  const int col = lane & 15, kr = lane >> 4;
  auto tile_of = [&](int t, int &I, int &J) {
    if (t < split) { I = I0; J = I0 + t; }
    else { I = I1; J = I1 + (t - split); }
  };
  auto compute_stage = [&](int b) {
#pragma unroll
    for (int k0 = 0; k0 < GBK; k0 += 4) {
      const double a0 = (I0 < nt) ? Ds[b][I0 * 16 + col][k0 + kr] : 0.0;
      const double a1 = (I1 < nt) ? Ds[b][I1 * 16 + col][k0 + kr] : 0.0;
#pragma unroll
      for (int t = 0; t < 17; t++) {
        int I, J;
        tile_of(t, I, J);
        if (I >= nt || J >= nt) continue;
        const double bv = Ds[b][J * 16 + col][k0 + kr];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(t < split ? a0 : a1, bv, acc[t], 0, 0, 0);
      }
    }
  };
  // stage k in LDS buffer k % 2; the registers carry stage k + 1 while
  // stage k is multiplied
  if (r0 < r1) load_stage(st0, r0);
  if (r0 < r1) store_stage(0, st0);
  if (r0 + GBK < r1) load_stage(st0, r0 + GBK);
  __syncthreads();
  for (int rs = r0; rs < r1; rs += 2 * GBK) {
    if (rs + GBK < r1) store_stage(1, st0);
    if (rs + 2 * GBK < r1) load_stage(st0, rs + 2 * GBK);
    compute_stage(0);
    __syncthreads();
    if (rs + GBK >= r1) break;
    if (rs + 2 * GBK < r1) store_stage(0, st0);
    if (rs + 3 * GBK < r1) load_stage(st0, rs + 3 * GBK);
    compute_stage(1);
    __syncthreads();
  }
  double *Pw = P + (size_t)blockIdx.x * n * n;
#pragma unroll
  for (int t = 0; t < 17; t++) {
    int I, J;
    tile_of(t, I, J);
    if (I >= nt || J >= nt) continue;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = I * 16 + kr + 4 * v, c = J * 16 + col;
      if (r < n && c < n) Pw[(size_t)r + (size_t)c * n] = acc[t][v];
    }
  }
}

// G(r, c) = sum over the partials in order, r <= c
__global__ __launch_bounds__(256) void gram_reduce_kernel(int n, int S, const double *__restrict__ P,
                                                          double *__restrict__ G, int ldg) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int r = e % n, c = e / n;
  if (r > c) return;
  double s = 0.0;
  for (int z = 0; z < S; z++) s += P[(size_t)z * n * n + e];
  G[(size_t)r + (size_t)c * ldg] = s;
}

}  // namespace

static int gram_launch(int f64, int m, int n, const double *A, int lda, double *G, int ldg) {
  if (n <= 0) return 0;
  if (n > GNB) return -1;
  if (m <= 0) {
    ld_laset('U', n, n, 0.0, 0.0, G, ldg);
    return 0;
  }
  hipStream_t st = (hipStream_t)ld_stream();
  // one workgroup per CU-sized share of the rows, chunks a multiple of the stage
  int S = 256;
  int chunk = (m + S - 1) / S;
  chunk = (chunk + GBK - 1) / GBK * GBK;
  S = (m + chunk - 1) / chunk;
  size_t mark = ld_ws_mark();
  double *P = (double *)ld_ws_alloc(sizeof(double) * (size_t)S * n * n);
  const int ph = ld_prof_begin(f64 ? "gram64_kernel" : "gram3_kernel");
  if (f64) hipLaunchKernelGGL(gram64_kernel, dim3(S), dim3(GNT), 0, st, m, n, A, lda, chunk, P);
  else hipLaunchKernelGGL(gram3_kernel, dim3(S), dim3(GNT), 0, st, m, n, A, lda, chunk, P);
  ld_prof_end(ph, (double)m * n * (n + 1), 8.0 * (double)m * n);
  const int tot = n * n;
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((tot + 255) / 256), dim3(256), 0, st, n, S, P, G, ldg);
  ld_ws_release(mark);
  return 0;
}

extern "C" int ld_gram_bf16x3(int m, int n, const double *A, int lda, double *G, int ldg) {
  if (ld_failed()) return 0;  // an earlier allocation of this call failed
  return gram_launch(0, m, n, A, lda, G, ldg);
}

extern "C" int ld_gram_f64(int m, int n, const double *A, int lda, double *G, int ldg) {
  if (ld_failed()) return 0;  // an earlier allocation of this call failed
  return gram_launch(1, m, n, A, lda, G, ldg);
}
