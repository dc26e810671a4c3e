// gram.hip -- the Gram matrix G = A^T A (upper triangle, n <= 256) of a
// tall FP64 matrix with 3-term split-BF16 MFMA products and FP32
// accumulation: each entry a = hi + mid + lo with three BF16 terms (the
// residuals a - hi and a - hi - mid are exact in FP64), and
// a^T b ~= hi^T hi' + hi^T mid' + mid^T hi' + hi^T lo' + lo^T hi' + mid^T mid'
// (the dropped terms are ~2^-24 |a||b|).  It is the first-pass Gram of the
// tall-skinny Cholesky-QR2 path (one read of A: HBM-bound at ~8 B per entry).
//
// Layout: one workgroup of 512 threads (8 waves) per row chunk; a stage of
// 32 rows x n columns is loaded into registers one stage ahead, split, and
// written to LDS as Ts[term][column][row] BF16 (row stride 40 halves: 16
// consecutive columns hit distinct banks); the 16 x 16 output tiles I <= J
// are dealt to the waves by row pairs (I = w and 15 - w) so each wave owns
// 17 tiles of a 16 x 16 tile grid and reuses its row operands across J.
// MFMA v_mfma_f32_16x16x16_bf16: lane l holds A(row l%16, k 4(l/16)..+3),
// B(k 4(l/16)..+3, col l%16) and D(row 4(l/16) + v, col l%16).  Partials
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
  __shared__ unsigned short Ts[3][GNB][GLD];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r0 = blockIdx.x * chunk;
  const int r1 = min(m, r0 + chunk);
  const int nt = (n + 15) / 16;
  // stage loads: thread t covers column t / 2, rows 16 (t % 2) .. + 15
  const int lc = tid >> 1, lh = (tid & 1) * 16;
  double st[16];
  auto load_stage = [&](int rs) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int r = rs + lh + q;
      st[q] = (lc < n && r < r1) ? A[(size_t)r + (size_t)lc * lda] : 0.0;
    }
  };
  auto store_stage = [&]() {
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
      *reinterpret_cast<bf4_t *>(&Ts[0][lc][lh + q4]) = t0;
      *reinterpret_cast<bf4_t *>(&Ts[1][lc][lh + q4]) = t1;
      *reinterpret_cast<bf4_t *>(&Ts[2][lc][lh + q4]) = t2;
    }
  };
  f4v_t acc[2][16];
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int j = 0; j < 16; j++) acc[s][j] = f4v_t{0.f, 0.f, 0.f, 0.f};
  const int rowt[2] = {w, 15 - w};
  const int col = lane & 15, kq = (lane >> 4) * 4;

  if (r0 < r1) load_stage(r0);
  for (int rs = r0; rs < r1; rs += GBK) {
    __syncthreads();  // the previous stage's LDS reads are done
    store_stage();
    __syncthreads();
    if (rs + GBK < r1) load_stage(rs + GBK);
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 16) {
#pragma unroll
      for (int s = 0; s < 2; s++) {
        const int I = rowt[s];
        if (I >= nt) continue;
        bf4_t a3[3];
#pragma unroll
        for (int t = 0; t < 3; t++)
          a3[t] = *reinterpret_cast<const bf4_t *>(&Ts[t][I * 16 + col][kk + kq]);
#pragma unroll
        for (int J = 0; J < 16; J++) {
          if (J < I || J >= nt) continue;
          bf4_t b3[3];
#pragma unroll
          for (int t = 0; t < 3; t++)
            b3[t] = *reinterpret_cast<const bf4_t *>(&Ts[t][J * 16 + col][kk + kq]);
          f4v_t c = acc[s][J];
          // small terms first
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3[1], b3[1], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3[0], b3[2], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3[2], b3[0], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3[0], b3[1], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3[1], b3[0], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3[0], b3[0], c, 0, 0, 0);
          acc[s][J] = c;
        }
      }
    }
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

extern "C" int ld_gram_bf16x3(int m, int n, const double *A, int lda, double *G, int ldg) {
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
  const int ph = ld_prof_begin("gram3_kernel");
  hipLaunchKernelGGL(gram3_kernel, dim3(S), dim3(GNT), 0, st, m, n, A, lda, chunk, P);
  ld_prof_end(ph, (double)m * n * (n + 1), 8.0 * (double)m * n);
  const int tot = n * n;
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((tot + 255) / 256), dim3(256), 0, st, n, S, P, G, ldg);
  ld_ws_release(mark);
  return 0;
}
