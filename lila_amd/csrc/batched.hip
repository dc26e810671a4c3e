// batched.hip -- many independent least-squares problems in one persistent
// launch (the LS sequence of least-squares/main_ls_qr3.c:94-119 per problem:
// Householder QR with dlarfg's convention, Q^T B through the block
// reflectors, back substitution with R).
//
// One workgroup (256 threads, 4 waves) owns one problem at a time and takes
// the next from a queue ordered by cost (largest first), so uneven sizes
// balance across the CUs.  Per problem, panels of NB = 16 columns:
//   1. the panel, column by column (one pass for the norm and the dots with
//      the later panel columns, one pass for the reflector update), on the
//      matrix in global memory;
//   2. W = V^T [V, A_trailing, B] on v_mfma_f64_16x16x4f64, the rows staged
//      through LDS 32 at a time with column-contiguous loads (the first 16
//      columns of W are the Gram V^T V the T factor needs);
//   3. T (dlarft 'F','C') from the Gram and tau; W2 = T^T W;
//   4. [A_trailing, B] -= V W2, again on MFMA with the tile staged in LDS.
// After the last panel, R X = (Q^T B)(0:n) is solved in place.
// Limits: m >= n, n <= 128, nrhs <= 16 (checked by the host).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include <algorithm>
#include "lila_dev.h"

typedef double d4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int BT = 256;              // threads per workgroup
constexpr int NB = 16;               // panel width
constexpr int RC = 32;               // rows per staged chunk
constexpr int MAXN = 128, MAXK = 16;
constexpr int MAXC = NB + MAXN + MAXK;  // 160 >= jb + trailing A columns + B columns
constexpr int LC = MAXC + 1;            // LDS row stride of the staged chunk (odd)

struct BProb {
  int m, n, k, lda, ldb, pad;
  double *A, *B;
};

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wmax(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

struct Smem {
  double C[RC][LC];         // staged rows: [V | A trailing | B], zero padded
  double W[NB][MAXC];       // V^T [V A B], then T^T W in the A/B part
  double T[NB][NB];
  double red[4][NB + 2];    // per-wave partial sums
  double tau[NB];
  double w[NB];             // tau * v^T a_q for the panel update
  double sc[4];             // beta, tau, 1/(alpha-beta)
  int job;
};

// element (r, q) of the staged column set: q < jb a reflector of the panel
// (unit lower structure), then the trailing A columns, then B's columns
__device__ __forceinline__ double col_val(const BProb &p, int j0, int jb, int na, int r, int q) {
  if (q < jb) {
    const int c = j0 + q;
    if (r < c) return 0.0;
    if (r == c) return 1.0;
    return p.A[(size_t)r + (size_t)c * p.lda];
  }
  q -= jb;
  if (q < na) return p.A[(size_t)r + (size_t)(j0 + jb + q) * p.lda];
  q -= na;
  return p.B[(size_t)r + (size_t)q * p.ldb];
}

__device__ void solve_problem(const BProb &p, Smem &s) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = p.m, n = p.n, k = p.k, lda = p.lda;
  double *A = p.A;
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int jb = min(NB, n - j0);
    // ---------------- 1. panel, column by column ----------------
    for (int c = 0; c < jb; c++) {
      const int col = j0 + c, nq = jb - c - 1;
      double ss = 0.0, mx = 0.0, d[NB - 1];
#pragma unroll
      for (int q = 0; q < NB - 1; q++) d[q] = 0.0;
      for (int r = col + 1 + tid; r < m; r += BT) {
        const double x = A[(size_t)r + (size_t)col * lda];
        ss += x * x;
        mx = fmax(mx, fabs(x));
#pragma unroll
        for (int q = 0; q < NB - 1; q++)
          if (q < nq) d[q] += x * A[(size_t)r + (size_t)(col + 1 + q) * lda];
      }
      ss = wsum(ss);
      mx = wmax(mx);
#pragma unroll
      for (int q = 0; q < NB - 1; q++) d[q] = wsum(d[q]);
      if (lane == 0) {
        s.red[wave][0] = ss;
        s.red[wave][1] = mx;
#pragma unroll
        for (int q = 0; q < NB - 1; q++) s.red[wave][2 + q] = d[q];
      }
      __syncthreads();
      if (tid == 0) {
        double sst = 0.0, mxt = 0.0;
        for (int w = 0; w < 4; w++) {
          sst += s.red[w][0];
          mxt = fmax(mxt, s.red[w][1]);
        }
        s.sc[3] = mxt;
        s.sc[2] = sst;
      }
      __syncthreads();
      double xnorm;
      const double mxt = s.sc[3];
      if (mxt != 0.0 && (mxt < 1e-150 || mxt > 1e150)) {
        // sum of squares out of range: a scaled pass (LAPACK dnrm2's purpose)
        const double inv = 1.0 / mxt;
        double s2 = 0.0;
        for (int r = col + 1 + tid; r < m; r += BT) {
          const double x = A[(size_t)r + (size_t)col * lda] * inv;
          s2 += x * x;
        }
        s2 = wsum(s2);
        __syncthreads();
        if (lane == 0) s.red[wave][0] = s2;
        __syncthreads();
        xnorm = mxt * sqrt(s.red[0][0] + s.red[1][0] + s.red[2][0] + s.red[3][0]);
      } else {
        xnorm = sqrt(s.sc[2]);
      }
      if (tid == 0) {
        const double alpha = A[(size_t)col + (size_t)col * lda];
        double beta = alpha, tau = 0.0, scale = 0.0;
        if (xnorm != 0.0) {
          const double aa = fabs(alpha);
          const double w = fmax(aa, xnorm), z = fmin(aa, xnorm);
          const double nrm = (z == 0.0) ? w : w * sqrt(1.0 + (z / w) * (z / w));
          beta = -copysign(nrm, alpha);
          tau = (beta - alpha) / beta;
          scale = 1.0 / (alpha - beta);
        }
        s.sc[0] = beta;
        s.sc[1] = tau;
        s.sc[2] = scale;
        s.tau[c] = tau;
        for (int q = 0; q < nq; q++) {
          double dq = 0.0;
          for (int w = 0; w < 4; w++) dq += s.red[w][2 + q];
          s.w[q] = tau * (A[(size_t)col + (size_t)(col + 1 + q) * lda] + dq * scale);
        }
      }
      __syncthreads();
      const double beta = s.sc[0], tau = s.sc[1], scale = s.sc[2];
      if (tau != 0.0) {
        for (int r = col + tid; r < m; r += BT) {
          if (r == col) {
            A[(size_t)col + (size_t)col * lda] = beta;
            for (int q = 0; q < nq; q++) A[(size_t)col + (size_t)(col + 1 + q) * lda] -= s.w[q];
          } else {
            const double v = A[(size_t)r + (size_t)col * lda] * scale;
            A[(size_t)r + (size_t)col * lda] = v;
            for (int q = 0; q < nq; q++) A[(size_t)r + (size_t)(col + 1 + q) * lda] -= v * s.w[q];
          }
        }
      }
      __syncthreads();
    }
    // ---------------- 2. W = V^T [V, A trailing, B] ----------------
    const int na = n - j0 - jb, ncf = jb + na + k;
    const int nct = (ncf + 15) / 16;
    d4_t acc[3];
#pragma unroll
    for (int u = 0; u < 3; u++) acc[u] = d4_t{0.0, 0.0, 0.0, 0.0};
    for (int rb = j0; rb < m; rb += RC) {
      // stage rows rb .. rb+RC-1 of all ncf columns (zero padded)
      for (int e = tid; e < RC * nct * 16; e += BT) {
        const int i = e % RC, q = e / RC, r = rb + i;
        s.C[i][q] = (r < m && q < ncf) ? col_val(p, j0, jb, na, r, q) : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 3; u++) {
        const int ct = wave + 4 * u;
        if (ct < nct) {
#pragma unroll
          for (int ks = 0; ks < RC / 4; ks++) {
            const int i = 4 * ks + (lane >> 4);
            const double va = s.C[i][lane & 15];            // V^T(p = lane%16, row i)
            const double cb = s.C[i][ct * 16 + (lane & 15)];  // C(row i, q)
            acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(va, cb, acc[u], 0, 0, 0);
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 3; u++) {
      const int ct = wave + 4 * u;
      if (ct < nct) {
#pragma unroll
        for (int v = 0; v < 4; v++) s.W[(lane >> 4) + 4 * v][ct * 16 + (lane & 15)] = acc[u][v];
      }
    }
    __syncthreads();
    // ---------------- 3. T from the Gram and tau; W2 = T^T W ----------------
    if (tid == 0) {
      for (int i = 0; i < jb; i++) {
        const double ti = s.tau[i];
        for (int r = 0; r < i; r++) {
          double acc2 = 0.0;
          if (ti != 0.0)
            for (int q = r; q < i; q++) acc2 += s.T[r][q] * s.W[q][i];
          s.T[r][i] = -ti * acc2;
        }
        s.T[i][i] = ti;
        for (int r = i + 1; r < NB; r++) s.T[r][i] = 0.0;
      }
    }
    __syncthreads();
    const int nc = na + k;
    // each thread owns one column of W: W2(:, q) = T^T W(:, q), rows >= jb zero
    for (int q = jb + tid; q < jb + nc; q += BT) {
      double wv[NB];
#pragma unroll
      for (int r = 0; r < NB; r++) wv[r] = s.W[r][q];
#pragma unroll
      for (int pr = 0; pr < NB; pr++) {
        double a2 = 0.0;
#pragma unroll
        for (int r = 0; r <= pr; r++) a2 += s.T[r][pr] * wv[r];
        s.W[pr][q] = (pr < jb) ? a2 : 0.0;
      }
    }
    __syncthreads();
    // ---------------- 4. [A trailing, B] -= V W2 ----------------
    if (nc > 0) {
      const int nct2 = (nc + 15) / 16;  // output column tiles over the A/B part
      for (int rb = j0; rb < m; rb += RC) {
        for (int e = tid; e < RC * (jb + nct2 * 16); e += BT) {
          const int i = e % RC, q = e / RC, r = rb + i;
          s.C[i][q] = (r < m && q < jb + nc) ? col_val(p, j0, jb, na, r, q) : 0.0;
        }
        __syncthreads();
        // tiles: 2 row blocks x nct2 column tiles, 4 MFMAs each (K = 16)
        for (int t = wave; t < 2 * nct2; t += 4) {
          const int rblk = t & 1, ct = t >> 1;
          d4_t dacc = d4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int ks = 0; ks < NB / 4; ks++) {
            const int pr = 4 * ks + (lane >> 4);
            const double va = s.C[rblk * 16 + (lane & 15)][pr];     // V(row, p)
            const double wb = s.W[pr][jb + ct * 16 + (lane & 15)];  // W2(p, q)
            dacc = __builtin_amdgcn_mfma_f64_16x16x4f64(va, wb, dacc, 0, 0, 0);
          }
#pragma unroll
          for (int v = 0; v < 4; v++) s.C[rblk * 16 + (lane >> 4) + 4 * v][jb + ct * 16 + (lane & 15)] -= dacc[v];
        }
        __syncthreads();
        // write the updated rows back, column-contiguous
        for (int e = tid; e < RC * nc; e += BT) {
          const int i = e % RC, q = e / RC, r = rb + i;
          if (r >= m) continue;
          const double v = s.C[i][jb + q];
          if (q < na) A[(size_t)r + (size_t)(j0 + jb + q) * lda] = v;
          else p.B[(size_t)r + (size_t)(q - na) * p.ldb] = v;
        }
        __syncthreads();
      }
    }
  }
  // ---------------- back substitution R X = (Q^T B)(0:n) ----------------
  for (int i = n - 1; i >= 0; i--) {
    const double rii = A[(size_t)i + (size_t)i * lda];
    if (tid < k) s.w[tid] = p.B[(size_t)i + (size_t)tid * p.ldb] / rii;
    __syncthreads();
    if (tid < k) p.B[(size_t)i + (size_t)tid * p.ldb] = s.w[tid];
    for (int e = tid; e < i * k; e += BT) {
      const int r = e % i, c = e / i;
      p.B[(size_t)r + (size_t)c * p.ldb] -= A[(size_t)r + (size_t)i * lda] * s.w[c];
    }
    __syncthreads();
  }
}

__global__ void batch_reset_kernel(int *counter) { *counter = 0; }

__global__ __launch_bounds__(BT) void batched_ls_kernel(const BProb *__restrict__ probs,
                                                       const int *__restrict__ order, int nprob,
                                                       int *counter) {
  __shared__ Smem s;
  for (;;) {
    if (threadIdx.x == 0) s.job = atomicAdd(counter, 1);
    __syncthreads();
    const int j = s.job;
    __syncthreads();
    if (j >= nprob) break;
    const BProb p = probs[order[j]];
    solve_problem(p, s);
    __syncthreads();
  }
}

}  // namespace

struct ld_batch_plan {
  int nprob, grid;
  BProb *dprob;
  int *dorder, *dcounter;
  double flops;
  hipGraph_t graph;
  hipGraphExec_t exec;
  int have_graph;
};

// flops of one problem, the count least-squares/main_ls_qr3.c:125-129 reports
static double prob_flops(double m, double n, double k) {
  return 2.0 * m * n * n - n * n * n / 3.0 + 4.0 * (m - n) * n * k + 4.0 * n * n * k;
}

extern "C" ld_batch_plan *ld_batch_plan_create(int nprob, const int *m, const int *n, const int *k,
                                               double *const *A, const int *lda, double *const *B,
                                               const int *ldb) {
  if (nprob <= 0) return nullptr;
  std::vector<BProb> hp(nprob);
  std::vector<double> cost(nprob);
  double fl = 0.0;
  for (int i = 0; i < nprob; i++) {
    if (m[i] < n[i] || n[i] < 1 || n[i] > MAXN || k[i] < 0 || k[i] > MAXK || lda[i] < m[i] ||
        (k[i] > 0 && ldb[i] < m[i])) {
      fprintf(stderr, "lila_hip: batched problem %d out of range (m=%d n=%d k=%d)\n", i, m[i], n[i],
              k[i]);
      return nullptr;
    }
    hp[i] = BProb{m[i], n[i], k[i], lda[i], ldb[i], 0, A[i], B[i]};
    cost[i] = prob_flops(m[i], n[i], k[i]);
    fl += cost[i];
  }
  std::vector<int> order(nprob);
  for (int i = 0; i < nprob; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  ld_batch_plan *pl = (ld_batch_plan *)calloc(1, sizeof(ld_batch_plan));
  pl->nprob = nprob;
  pl->flops = fl;
  pl->dprob = (BProb *)ld_dev_malloc(sizeof(BProb) * nprob);
  pl->dorder = (int *)ld_dev_malloc(sizeof(int) * nprob);
  pl->dcounter = (int *)ld_dev_malloc(sizeof(int));
  if (!pl->dprob || !pl->dorder || !pl->dcounter) {
    free(pl);
    return nullptr;
  }
  ld_memcpy_h2d(pl->dprob, hp.data(), sizeof(BProb) * nprob);
  ld_memcpy_h2d(pl->dorder, order.data(), sizeof(int) * nprob);
  int dev = 0, ncu = 0, per = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)batched_ls_kernel, BT, 0);
  if (per < 1) per = 1;
  pl->grid = ncu * per;
  if (pl->grid > nprob) pl->grid = nprob;
  ld_sync();
  return pl;
}

// enqueue: reset the queue counter, then the persistent kernel
static void batch_enqueue(ld_batch_plan *pl, hipStream_t st) {
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(1), 0, st, pl->dcounter);
  hipLaunchKernelGGL(batched_ls_kernel, dim3(pl->grid), dim3(BT), 0, st, pl->dprob, pl->dorder,
                     pl->nprob, pl->dcounter);
}

extern "C" int ld_batch_plan_run(ld_batch_plan *pl, int use_graph) {
  if (!pl) return -1;
  hipStream_t st = (hipStream_t)ld_stream();
  const int ph = ld_prof_begin("batched_ls_kernel");
  if (use_graph) {
    if (!pl->have_graph) {
      // capture [reset, persistent kernel] once; replays launch both as one graph
      hipStream_t cs;
      if (hipStreamCreate(&cs) != hipSuccess) return -1;
      if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) return -1;
      batch_enqueue(pl, cs);
      if (hipStreamEndCapture(cs, &pl->graph) != hipSuccess) return -1;
      if (hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0) != hipSuccess) return -1;
      hipStreamDestroy(cs);
      pl->have_graph = 1;
    }
    if (hipGraphLaunch(pl->exec, st) != hipSuccess) return -1;
  } else {
    batch_enqueue(pl, st);
  }
  ld_prof_end(ph, pl->flops, 0.0);
  return 0;
}

extern "C" double ld_batch_plan_flops(const ld_batch_plan *pl) { return pl ? pl->flops : 0.0; }

extern "C" void ld_batch_plan_destroy(ld_batch_plan *pl) {
  if (!pl) return;
  ld_sync();
  if (pl->have_graph) {
    hipGraphExecDestroy(pl->exec);
    hipGraphDestroy(pl->graph);
  }
  ld_dev_free(pl->dprob);
  ld_dev_free(pl->dorder);
  ld_dev_free(pl->dcounter);
  free(pl);
}
