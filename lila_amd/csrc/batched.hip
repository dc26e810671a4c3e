// batched.hip -- many independent least-squares problems in one persistent
// launch: per problem the sequence of least-squares/main_ls_qr3.c:94-119
// (Householder QR with dlarfg's convention, Q^T B through the block
// reflectors, back substitution with R), outputs in LAPACK dgels' layout.
//
// One workgroup (512 threads, 8 waves) owns one problem at a time and takes
// the next from a queue ordered by flop count (largest first), so uneven
// sizes balance across the CUs.  A problem is factored in panels of 32
// columns; the trailing columns of A and the columns of B are carried
// together ("trailing" below).  Every pass streams the rows below the
// panel's top in chunks of 64 rows through LDS (one wave reads 512
// contiguous bytes of a column; two chunks prefetched into registers while
// the current one is used; barriers inside a pass order LDS only, so the
// prefetches stay in flight) and does its arithmetic on
// v_mfma_f64_16x16x4f64:
//   a. G1 = P^T P for the first panel P (later panels get G1 from pass U);
//   b. Cholesky G1 = R1^T R1 (one wave), Q1 = P R1^-1 (R1^-1 formed
//      explicitly) and G2 = Q1^T Q1;
//      Cholesky G2 = R2^T R2, R = R2 R1, the top block of Q = Q1 R2^-1;
//      Householder reconstruction: sign-chosen LU Q_top S - I = L U,
//      V1 = L, R_hh = S R, T = -U V1^-T;
//   d. V2 = P2 (R^-1 S U^-1) written in place, and W = V^T [V | trailing]
//      accumulated in the same pass (V^T V too when the panel took the
//      column loop); W2 = T^T W;
//   U. trailing -= V W2, and G1 of the next panel accumulated from the
//      updated columns in the same pass.
// So each panel costs two reads of the panel and two reads plus one write
// of the trailing block.  If a Cholesky pivot is not safely positive (or
// R's diagonal spreads by more than 1e5, or a reflector has tau within
// 1e-10 of 0 or 2) the panel is factored column by column instead (dlarfg
// and the reflector update on the panel in global memory) and T comes from
// dlarft over the Gram V^T V of pass d.  After the last panel R X =
// (Q^T B)(0:n) is solved in blocks of 32 rows.
// Limits: m >= n >= 1, n <= 128, 0 <= nrhs <= 16 (checked by the host).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include <algorithm>
#include "lila_dev.h"
#include "small32.h"

typedef double d4_t __attribute__((ext_vector_type(4)));

namespace {

using small32::LDM;
using small32::PNB;

constexpr int BT = 512;                 // threads per workgroup
constexpr int NW = BT / 64;             // waves
constexpr int NB = 32;                  // panel width
constexpr int RC = 64;                  // rows per staged chunk
constexpr int MAXN = 128, MAXK = 16;
constexpr int MAXT = MAXN - NB + MAXK;  // 112 trailing columns at most
constexpr int CW = NB + MAXT;           // 144 chunk columns
constexpr int LC = CW + 1;              // LDS row stride of the chunk (odd)
constexpr int QS = BT / RC;             // 8: column step between a thread's elements
constexpr int PF = RC * CW / BT;        // 18 staged elements per thread
constexpr int UP = NB / QS;             // 4 of them in the panel columns
static_assert(NB == PNB && RC * CW % BT == 0 && NB % QS == 0 && (RC / 16) * 2 == NW, "chunk mapping");

struct BProb {
  int m, n, k, lda, ldb, pad;
  double *A, *B;
};

struct Smem {
  double C[RC][LC];     // chunk rows: panel columns 0..31, trailing columns from 32
  double W[NB][CW];     // V^T [V | trailing]; the trailing part becomes T^T W
  double M0[PNB][LDM];  // R1, then R = R2 R1
  double M2[PNB][LDM];  // R2
  double M3[PNB][LDM];  // top block of Q, then its LU (U upper, L strict lower)
  double MX[PNB][LDM];  // transform of a row pass: R1^-1, then R^-1 S U^-1
  double TT[PNB][LDM];  // R2^-1 / U^-1 scratch, then T
  double D0[PNB], D2[PNB], D3[PNB], Sv[PNB], tau[PNB], wv[PNB];
  double red[NW][PNB + 2];
  double sc[4];
  int job, flag, colbad, near2;
};

// the panel being factored: first column j0, width jb, na A columns right
// of it, nt = na + k trailing columns
struct Pan {
  int j0, jb, na, nt;
};

// thread index the compiler cannot hoist (address and mask computations
// derived from it stay next to their use instead of living, spilled, for the
// whole persistent kernel)
__device__ __forceinline__ int otid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

#define THREAD_IDS                                   \
  const int tid = otid(), lane = tid & 63, wave = tid >> 6; \
  (void)lane;                                        \
  (void)wave

// workgroup barrier ordering LDS only: the prefetch loads of the next chunk
// stay in flight across it (__syncthreads() would wait for every
// outstanding global access, prefetches included)
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ d4_t mfma(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

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

// ---- chunk staging: thread element u is row tid % RC of the chunk and
// chunk column tid / RC + QS u (u < UP: panel columns, then trailing) ----
__device__ __forceinline__ void fetch(const BProb &p, const Pan &P, int rb, bool trail,
                                      double (&reg)[PF]) {
  const int tid = otid(), i = tid % RC, q0 = tid / RC, r = rb + i;
  const bool rok = r < p.m;
#pragma unroll
  for (int u = 0; u < PF; u++) {
    const int q = q0 + QS * u;
    bool ok;
    const double *src;
    if (u < UP) {
      ok = q < P.jb;
      src = p.A + (size_t)(P.j0 + q) * p.lda;
    } else {
      const int t = q - NB;
      ok = trail && t < P.nt;
      src = (t < P.na) ? p.A + (size_t)(P.j0 + P.jb + t) * p.lda
                       : p.B + (size_t)(t - P.na) * p.ldb;
    }
    reg[u] = (rok && ok) ? src[r] : 0.0;
  }
}

// vmask: the chunk holds the panel's top rows and its panel columns are V,
// stored with R above the diagonal: give them V1's unit lower structure
__device__ __forceinline__ void put(Smem &s, const double (&reg)[PF], bool vmask, int jb) {
  const int tid = otid(), i = tid % RC, q0 = tid / RC;
#pragma unroll
  for (int u = 0; u < PF; u++) {
    const int q = q0 + QS * u;
    double v = reg[u];
    if (u < UP && vmask) v = (i < q) ? 0.0 : (i == q ? (q < jb ? 1.0 : 0.0) : v);
    s.C[i][q] = v;
  }
}

__device__ __forceinline__ void store_trail(const Smem &s, const BProb &p, const Pan &P, int rb) {
  const int tid = otid(), i = tid % RC, q0 = tid / RC, r = rb + i;
  if (r >= p.m) return;
#pragma unroll
  for (int u = UP; u < PF; u++) {
    const int t = q0 + QS * u - NB;
    if (t < P.nt) {
      double *dst = (t < P.na) ? p.A + (size_t)(P.j0 + P.jb + t) * p.lda
                               : p.B + (size_t)(t - P.na) * p.ldb;
      dst[r] = s.C[i][NB + t];
    }
  }
}

// rows rlo.. of the chunk
__device__ __forceinline__ void store_panel(const Smem &s, const BProb &p, const Pan &P, int rb,
                                            int rlo) {
  const int tid = otid(), i = tid % RC, q0 = tid / RC, r = rb + i;
  if (r >= p.m || i < rlo) return;
#pragma unroll
  for (int u = 0; u < UP; u++) {
    const int q = q0 + QS * u;
    if (q < P.jb) p.A[(size_t)r + (size_t)(P.j0 + q) * p.lda] = s.C[i][q];
  }
}

// Stream chunks r0, r0 + RC, ... (nch of them) through LDS: two chunks are
// prefetched into registers while body(ch) works on the one in LDS.  vmask0:
// the first chunk's panel columns take V1's structure (see put).
template <typename Body>
__device__ __forceinline__ void stream(Smem &s, const BProb &p, const Pan &P, int r0, int nch,
                                       bool trail, bool vmask0, Body &&body) {
  double ra[PF], rb[PF];
  fetch(p, P, r0, trail, ra);
  if (nch > 1) fetch(p, P, r0 + RC, trail, rb);
  for (int ch = 0; ch < nch; ch += 2) {
    put(s, ra, vmask0 && ch == 0, P.jb);
    lds_barrier();
    if (ch + 2 < nch) fetch(p, P, r0 + (ch + 2) * RC, trail, ra);
    body(ch);
    lds_barrier();
    if (ch + 1 < nch) {
      put(s, rb, false, P.jb);
      lds_barrier();
      if (ch + 3 < nch) fetch(p, P, r0 + (ch + 3) * RC, trail, rb);
      body(ch + 1);
      lds_barrier();
    }
  }
  __syncthreads();  // the pass's global stores visible to the workgroup
}

// ---- 32 x 32 pieces ----
// Gram tile tt (0: (0,0), 1: (0,1), 2: (1,1)) of X^T X, X = chunk columns
// c0..c0+ncol-1 (zero beyond), over the chunk rows from ilo (a multiple of 4) on
__device__ __forceinline__ d4_t gram_tile(const Smem &s, int c0, int tt, d4_t acc, int ilo = 0,
                                          int ncol = NB) {
  const int lane = otid() & 63;
  const int ti = (tt == 2) ? 1 : 0, tj = (tt == 0) ? 0 : 1;
  // columns from ncol on are left out (they hold other data)
  const bool va = ti * 16 + (lane & 15) < ncol, vb = tj * 16 + (lane & 15) < ncol;
#pragma unroll
  for (int ks = 0; ks < RC / 4; ks++) {
    const int i = 4 * ks + (lane >> 4);
    if (4 * ks >= ilo)
      acc = mfma(va ? s.C[i][c0 + ti * 16 + (lane & 15)] : 0.0,
                 vb ? s.C[i][c0 + tj * 16 + (lane & 15)] : 0.0, acc);
  }
  return acc;
}
__device__ __forceinline__ void gram_store(double (*M)[LDM], int tt, d4_t acc) {
  const int lane = otid() & 63;
  const int ti = (tt == 2) ? 1 : 0, tj = (tt == 0) ? 0 : 1;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    M[ti * 16 + (lane >> 4) + 4 * v][tj * 16 + (lane & 15)] = acc[v];
    if (tt == 1) M[16 + (lane >> 4) + 4 * v][lane & 15] = 0.0;
  }
}
// identity beyond jb on the diagonal (the data there is zero)
__device__ __forceinline__ void pad_identity(double (*M)[LDM], int jb) {
  const int tid = otid();
  if (tid < PNB && tid >= jb) M[tid][tid] = 1.0;
}

// (X Y)(r, c) for upper X, Y; colscale multiplies X's column q by sc[q]
__device__ __forceinline__ double upper_prod(const double (*X)[LDM], const double (*Y)[LDM],
                                             const double *sc, int r, int c) {
  double acc = 0.0;
  for (int q = r; q <= c; q++) acc += X[r][q] * (sc ? sc[q] : 1.0) * Y[q][c];
  return acc;
}

// ---- column-by-column fallback: dlarfg and the reflector update on the
// panel in global memory; tau into s.tau ----
// (scalar arguments: a struct passed by reference to an outlined call would
// be kept in scratch memory by the caller)
__device__ __noinline__ void panel_columns(double *A, int m, int lda, int j0, int jb, Smem &s) {
  const int tid = otid(), lane = tid & 63, wave = tid >> 6;
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
      int bad = 0;
      for (int w = 0; w < NW; w++) {
        sst += s.red[w][0];
        mxt = fmax(mxt, s.red[w][1]);
      }
      for (int q = 0; q < nq; q++) {
        double dq = 0.0;
        for (int w = 0; w < NW; w++) dq += s.red[w][2 + q];
        s.wv[q] = dq;
        bad |= !(fabs(dq) <= 1.7976931348623157e308);
      }
      s.sc[3] = mxt;
      s.sc[2] = sst;
      s.colbad = bad;
    }
    __syncthreads();
    double xnorm, sc = 1.0;
    const double mxt = s.sc[3];
    if (mxt != 0.0 && (mxt < 1e-150 || mxt > 1e150 || s.colbad)) {
      // a sum of squares or a product out of range: a pass on the column
      // scaled by a power of two near 1 / max|x| (the purpose of dnrm2's
      // scaling), the dot products with the later columns included
      const double inv = ldexp(1.0, -ilogb(mxt));
      sc = ldexp(1.0, ilogb(mxt));
      double s2 = 0.0;
#pragma unroll
      for (int q = 0; q < NB - 1; q++) d[q] = 0.0;
      for (int r = col + 1 + tid; r < m; r += BT) {
        const double x = A[(size_t)r + (size_t)col * lda] * inv;
        s2 += x * x;
#pragma unroll
        for (int q = 0; q < NB - 1; q++)
          if (q < nq) d[q] += x * A[(size_t)r + (size_t)(col + 1 + q) * lda];
      }
      s2 = wsum(s2);
#pragma unroll
      for (int q = 0; q < NB - 1; q++) d[q] = wsum(d[q]);
      __syncthreads();
      if (lane == 0) {
        s.red[wave][0] = s2;
#pragma unroll
        for (int q = 0; q < NB - 1; q++) s.red[wave][2 + q] = d[q];
      }
      __syncthreads();
      double t = 0.0;
      for (int w = 0; w < NW; w++) t += s.red[w][0];
      xnorm = sc * sqrt(t);
      if (tid == 0)
        for (int q = 0; q < nq; q++) {
          double dq = 0.0;
          for (int w = 0; w < NW; w++) dq += s.red[w][2 + q];
          s.wv[q] = dq;
        }
    } else {
      xnorm = sqrt(s.sc[2]);
    }
    if (tid == 0) {
      const double alpha = A[(size_t)col + (size_t)col * lda];
      double beta = alpha, tau = 0.0, scale = 0.0, dscale = 0.0;
      if (xnorm != 0.0) {
        const double aa = fabs(alpha);
        const double w = fmax(aa, xnorm), z = fmin(aa, xnorm);
        const double nrm = (z == 0.0) ? w : w * sqrt(1.0 + (z / w) * (z / w));
        beta = -copysign(nrm, alpha);
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
        dscale = sc / (alpha - beta);  // the dots are of the column scaled by 1 / sc
      }
      s.sc[0] = beta;
      s.sc[1] = tau;
      s.sc[2] = scale;
      s.tau[c] = tau;
      for (int q = 0; q < nq; q++)
        s.wv[q] = tau * (A[(size_t)col + (size_t)(col + 1 + q) * lda] + s.wv[q] * dscale);
    }
    __syncthreads();
    const double beta = s.sc[0], tau = s.sc[1], scale = s.sc[2];
    if (tau != 0.0) {
      for (int r = col + tid; r < m; r += BT) {
        if (r == col) {
          A[(size_t)col + (size_t)col * lda] = beta;
          for (int q = 0; q < nq; q++) A[(size_t)col + (size_t)(col + 1 + q) * lda] -= s.wv[q];
        } else {
          const double v = A[(size_t)r + (size_t)col * lda] * scale;
          A[(size_t)r + (size_t)col * lda] = v;
          for (int q = 0; q < nq; q++) A[(size_t)r + (size_t)(col + 1 + q) * lda] -= v * s.wv[q];
        }
      }
    }
    __syncthreads();
  }
}

// ---- one least-squares problem ----
__device__ void solve_problem(const BProb &p, Smem &s) {
  const int m = p.m, n = p.n, k = p.k;
  // pass a: G1 of the first panel
  {
    THREAD_IDS;
    const Pan P{0, min(NB, n), n - min(NB, n), n - min(NB, n) + k};
    d4_t g = {0.0, 0.0, 0.0, 0.0};
    const int nch = (m + RC - 1) / RC;
    stream(s, p, P, 0, nch, false, false, [&](int ch) {
      THREAD_IDS;
      if (wave < 3) g = gram_tile(s, 0, wave, g);
    });
    if (wave < 3) gram_store(s.M0, wave, g);
    __syncthreads();
    pad_identity(s.M0, P.jb);
    __syncthreads();
  }

  for (int j0 = 0; j0 < n; j0 += NB) {
    THREAD_IDS;
    const int jb = min(NB, n - j0), na = n - j0 - jb;
    const Pan P{j0, jb, na, na + k};
    const int nch = (m - j0 + RC - 1) / RC;
    const int nct = (P.nt + 15) / 16;  // column tiles of the trailing block

    // ---- Cholesky-QR2 + reconstruction of the panel ----
    if (wave == 0) {
      const bool ok = small32::wave_chol32(&s.M0[0][0], s.D0, jb);
      if (tid == 0) s.flag = ok ? 1 : 0;
    }
    __syncthreads();
    bool cqr = s.flag != 0, near2 = false;
    if (cqr) {
      if (wave == 0) small32::inv_upper32(&s.M0[0][0], s.D0, &s.MX[0][0]);
      __syncthreads();
      // pass b: Q1 = P R1^-1 (into chunk columns 32..63), G2 = Q1^T Q1, and
      // the top block of Q1 into M3
      d4_t g = {0.0, 0.0, 0.0, 0.0};
      stream(s, p, P, j0, nch, false, false, [&](int ch) {
        THREAD_IDS;
        {
          // RC / 16 x 2 tiles, one per wave
          const int ri = wave % (RC / 16), cj = wave / (RC / 16);
          d4_t q = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int ks = 0; ks < NB / 4; ks++) {
            const int c = 4 * ks + (lane >> 4);
            if (cj == 1 || ks < NB / 8)  // MX is upper triangular
              q = mfma(s.C[ri * 16 + (lane & 15)][c], s.MX[c][cj * 16 + (lane & 15)], q);
          }
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int r = ri * 16 + (lane >> 4) + 4 * v, c = cj * 16 + (lane & 15);
            s.C[r][NB + c] = q[v];
            if (ch == 0 && r < PNB) s.M3[r][c] = q[v];
          }
        }
        lds_barrier();
        if (wave >= 4 && wave < 7) g = gram_tile(s, NB, wave - 4, g);
      });
      if (wave >= 4 && wave < 7) gram_store(s.M2, wave - 4, g);
      __syncthreads();
      pad_identity(s.M2, jb);
      if (tid == 0) s.near2 = 1;
      __syncthreads();
      // G2 within 1e-11 of I entrywise (the usual case after a first pass
      // that worked): R2 = I + X, R2^-1 = I - X (X: G2 - I above the
      // diagonal, half of it on it; residuals below 1e-18), no Cholesky chain
      for (int e = tid; e < PNB * PNB; e += BT) {
        const int r = e / PNB, c = e % PNB;
        if (r <= c && !(fabs(r == c ? s.M2[r][c] - 1.0 : s.M2[r][c]) <= 1e-11)) s.near2 = 0;
      }
      __syncthreads();
      near2 = s.near2 != 0;
      if (near2) {
        for (int e = tid; e < PNB * PNB; e += BT) {
          const int r = e / PNB, c = e % PNB;
          const double gv = s.M2[r][c];
          s.M2[r][c] = (r < c) ? gv : (r == c ? 0.5 * (1.0 + gv) : 0.0);
          s.TT[r][c] = (r < c) ? -gv : (r == c ? 0.5 * (3.0 - gv) : 0.0);
        }
        if (tid == 0) s.flag = 1;
      } else if (wave == 0) {
        const bool ok = small32::wave_chol32(&s.M2[0][0], s.D2, jb);
        if (tid == 0) s.flag = ok ? 1 : 0;
      }
      __syncthreads();
      cqr = s.flag != 0;
    }
    if (cqr) {
      if (!near2) {
        if (wave == 0) small32::inv_upper32(&s.M2[0][0], s.D2, &s.TT[0][0]);  // R2^-1
        __syncthreads();
      }
      // Q_top = Q1_top R2^-1, R^-1 = R1^-1 R2^-1, R = R2 R1
      double qt[2], ri[2], rr[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int e = tid + h * BT, r = e / PNB, c = e % PNB;
        double acc = 0.0;
        for (int q = 0; q <= c; q++) acc += s.M3[r][q] * s.TT[q][c];
        qt[h] = acc;
        ri[h] = (r <= c) ? upper_prod(s.MX, s.TT, nullptr, r, c) : 0.0;
        rr[h] = (r <= c) ? upper_prod(s.M2, s.M0, nullptr, r, c) : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int e = tid + h * BT, r = e / PNB, c = e % PNB;
        s.M3[r][c] = (r < jb && c < jb) ? qt[h] : (r == c ? 1.0 : 0.0);
        s.MX[r][c] = ri[h];
        s.M0[r][c] = rr[h];
      }
      __syncthreads();
      if (wave == 0) {
        const bool ok = small32::wave_lu_sign32(&s.M3[0][0], s.Sv, s.D3, jb);
        if (tid == 0) s.flag = ok ? 1 : 0;
      }
      __syncthreads();
      cqr = s.flag != 0;
    }
    if (cqr) {
      if (wave == 0) small32::inv_upper32(&s.M3[0][0], s.D3, &s.TT[0][0]);  // U^-1 (M3's upper part)
      __syncthreads();
      // transform of pass d: R^-1 S U^-1
      double mf[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int e = tid + h * BT, r = e / PNB, c = e % PNB;
        mf[h] = (r <= c) ? upper_prod(s.MX, s.TT, s.Sv, r, c) : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int e = tid + h * BT;
        s.MX[e / PNB][e % PNB] = mf[h];
      }
      // T = -U V1^-T, row r by thread r: T V1^T = -U, right-looking over p
      if (tid < PNB) {
        const int r = tid;
        double t[PNB];
#pragma unroll
        for (int c = 0; c < PNB; c++) t[c] = (c >= r) ? -s.M3[r][c] : 0.0;
#pragma unroll
        for (int q = 0; q < PNB; q++) {
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int c = q + 1; c < PNB; c++) t[c] -= t[q] * s.M3[c][q];
        }
#pragma unroll
        for (int c = 0; c < PNB; c++) s.TT[r][c] = (r < jb && c < jb) ? t[c] : 0.0;
      }
      __syncthreads();
    } else {
      panel_columns(p.A, m, p.lda, j0, jb, s);
    }

    // ---- pass d: V (CQR: formed and written), W = V^T [V | trailing] ----
    d4_t wacc[3];
#pragma unroll
    for (int u = 0; u < 3; u++) wacc[u] = d4_t{0.0, 0.0, 0.0, 0.0};
    // W tiles (row tile ti, column tile tj); the Gram V^T V (tj < 2) only
    // for the column-loop panel's dlarft
    const int tj0 = cqr ? 2 : 0, ntiles = 2 * (2 + nct - tj0);
    stream(s, p, P, j0, nch, true, !cqr, [&](int ch) {
      THREAD_IDS;
      const int rb = j0 + ch * RC;
      if (cqr) {
        // V = P (R^-1 S U^-1) on every row of the chunk (RC / 16 x 2 tiles,
        // one per wave)
        d4_t q4 = {0.0, 0.0, 0.0, 0.0};
        const int ri = wave % (RC / 16), cj = wave / (RC / 16);
#pragma unroll
        for (int ks = 0; ks < NB / 4; ks++) {
          const int c = 4 * ks + (lane >> 4);
          if (cj == 1 || ks < NB / 8)  // MX is upper triangular
            q4 = mfma(s.C[ri * 16 + (lane & 15)][c], s.MX[c][cj * 16 + (lane & 15)], q4);
        }
        lds_barrier();
#pragma unroll
        for (int v = 0; v < 4; v++) s.C[ri * 16 + (lane >> 4) + 4 * v][cj * 16 + (lane & 15)] = q4[v];
        lds_barrier();
        if (ch == 0) {
          // the panel's top jb rows: V1 = L in the chunk; R_hh = S R on and
          // above the diagonal and L below it into A
          for (int e = tid; e < jb * NB; e += BT) {
            const int i = e % jb, q = e / jb;
            const double l = s.M3[i][q];
            s.C[i][q] = (q < jb && i >= q) ? (i == q ? 1.0 : l) : 0.0;
            if (q < jb)
              p.A[(size_t)(j0 + i) + (size_t)(j0 + q) * p.lda] = (i <= q) ? s.Sv[i] * s.M0[i][q] : l;
          }
        }
        store_panel(s, p, P, rb, ch == 0 ? jb : 0);
        lds_barrier();
      }
#pragma unroll
      for (int u = 0; u < 3; u++) {
        const int id = wave + NW * u;
        if (id < ntiles) {
          const int ti = id & 1, tj = tj0 + (id >> 1);
#pragma unroll
          for (int ks = 0; ks < RC / 4; ks++) {
            const int i = 4 * ks + (lane >> 4);
            wacc[u] = mfma(s.C[i][ti * 16 + (lane & 15)], s.C[i][tj * 16 + (lane & 15)], wacc[u]);
          }
        }
      }
    });
#pragma unroll
    for (int u = 0; u < 3; u++) {
      const int id = wave + NW * u;
      if (id < ntiles) {
        const int ti = id & 1, tj = tj0 + (id >> 1);
#pragma unroll
        for (int v = 0; v < 4; v++) s.W[ti * 16 + (lane >> 4) + 4 * v][tj * 16 + (lane & 15)] = wacc[u][v];
      }
    }
    __syncthreads();
    if (!cqr) {
      // dlarft('F','C') from the Gram V^T V (W's first 32 columns) and tau, one wave
      if (wave == 0) {
        for (int i = 0; i < PNB; i++) {
          if (lane < PNB) {
            const double ti = (i < jb) ? s.tau[i] : 0.0;
            double acc = 0.0;
            if (lane < i)
              for (int q = lane; q < i; q++) acc += s.TT[lane][q] * s.W[q][i];
            s.TT[lane][i] = (lane < i) ? -ti * acc : (lane == i ? ti : 0.0);
          }
          small32::wfence();
        }
      }
      __syncthreads();
    }
    if (P.nt == 0) continue;  // nothing right of the last panel and no B

    // ---- W2 = T^T W over the trailing columns ----
    {
      d4_t w2[2];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        w2[u] = d4_t{0.0, 0.0, 0.0, 0.0};
        const int id = wave + NW * u;
        if (id < 2 * nct) {
          const int ti = id & 1, tj = id >> 1;
#pragma unroll
          for (int ks = 0; ks < NB / 4; ks++) {
            const int q = 4 * ks + (lane >> 4);
            w2[u] = mfma(s.TT[q][ti * 16 + (lane & 15)], s.W[q][NB + tj * 16 + (lane & 15)], w2[u]);
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int id = wave + NW * u;
        if (id < 2 * nct) {
          const int ti = id & 1, tj = id >> 1;
#pragma unroll
          for (int v = 0; v < 4; v++) s.W[ti * 16 + (lane >> 4) + 4 * v][NB + tj * 16 + (lane & 15)] = w2[u][v];
        }
      }
      __syncthreads();
    }

    // ---- pass U: trailing -= V W2; G1 of the next panel ----
    const bool next = na > 0;
    const int jbn = min(NB, na);  // width of the next panel
    d4_t g = {0.0, 0.0, 0.0, 0.0};
    stream(s, p, P, j0, nch, true, true, [&](int ch) {
      THREAD_IDS;
      const int rb = j0 + ch * RC;
#pragma unroll
      for (int u = 0; u < (RC / 16) * (MAXT / 16) / NW + 1; u++) {
        const int id = wave + NW * u;
        if (id < (RC / 16) * nct) {
          const int ri = id % (RC / 16), tj = id / (RC / 16);
          d4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int ks = 0; ks < NB / 4; ks++) {
            const int q = 4 * ks + (lane >> 4);
            acc = mfma(s.C[ri * 16 + (lane & 15)][q], s.W[q][NB + tj * 16 + (lane & 15)], acc);
          }
#pragma unroll
          for (int v = 0; v < 4; v++) s.C[ri * 16 + (lane >> 4) + 4 * v][NB + tj * 16 + (lane & 15)] -= acc[v];
        }
      }
      lds_barrier();
      // rows from j0 + 32 on (the next panel's rows) and its jbn columns
      if (next && wave >= 5) g = gram_tile(s, NB, wave - 5, g, ch == 0 ? NB : 0, jbn);
      store_trail(s, p, P, rb);
    });
    if (next) {
      if (wave >= 5) gram_store(s.M0, wave - 5, g);
      __syncthreads();
      pad_identity(s.M0, jbn);
      __syncthreads();
    }
  }

  // ---- back substitution R X = (Q^T B)(0:n), blocks of 32 rows from the bottom ----
  if (k > 0) {
    for (int b0 = ((n - 1) / NB) * NB; b0 >= 0; b0 -= NB) {
      THREAD_IDS;
      const int bb = min(NB, n - b0);
      for (int e = tid; e < PNB * PNB; e += BT) {
        const int i = e % PNB, c = e / PNB;
        s.M0[i][c] = (i < bb && c < bb) ? (i <= c ? p.A[(size_t)(b0 + i) + (size_t)(b0 + c) * p.lda] : 0.0)
                                        : (i == c ? 1.0 : 0.0);
        if (c < k) s.C[i][c] = (i < bb) ? p.B[(size_t)(b0 + i) + (size_t)c * p.ldb] : 0.0;
      }
      __syncthreads();
      if (tid < k) {
        double y[PNB];
#pragma unroll
        for (int i = 0; i < PNB; i++) y[i] = s.C[i][tid];
#pragma unroll
        for (int i = PNB - 1; i >= 0; i--) {
          y[i] = y[i] / s.M0[i][i];
#pragma unroll
          for (int r = 0; r < i; r++) y[r] -= s.M0[r][i] * y[i];
        }
#pragma unroll
        for (int i = 0; i < PNB; i++) s.C[i][tid] = y[i];
      }
      __syncthreads();
      for (int e = tid; e < bb * k; e += BT) {
        const int i = e % bb, c = e / bb;
        p.B[(size_t)(b0 + i) + (size_t)c * p.ldb] = s.C[i][c];
      }
      for (int e = tid; e < b0 * k; e += BT) {
        const int r = e % b0, c = e / b0;
        double acc = 0.0;
        for (int q = 0; q < bb; q++) acc += p.A[(size_t)r + (size_t)(b0 + q) * p.lda] * s.C[q][c];
        p.B[(size_t)r + (size_t)c * p.ldb] -= acc;
      }
      __syncthreads();
    }
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
  const int cap = ld_tune()->batch_grid;
  if (cap > 0 && pl->grid > cap) pl->grid = cap;
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

// ---------------------------------------------------------------------------
// Batched least squares without allocation: the problem table, the queue
// order and the queue counter live in `scratch`, a file-scope device array
// (one copy per GPU, held until the process ends).  Host code never uses the
// array's name as an address: its device address is looked up once per
// device with hipGetSymbolAddress and kept in na_ptr[device].  Every call on
// a GPU shares that GPU's array; calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace {
constexpr int NA_MAXPROB = 16384;  // problems per call at most
constexpr size_t NA_TAB = sizeof(BProb) * NA_MAXPROB;
constexpr size_t NA_ORD = sizeof(int) * NA_MAXPROB;
constexpr size_t NA_BYTES = NA_TAB + NA_ORD + 256;  // + the queue counter
}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {
constexpr int NA_MAXDEV = 64;
void *na_ptr[NA_MAXDEV];
int na_grid[NA_MAXDEV];
// host staging of the table (a pageable source: hipMemcpyAsync copies it
// synchronously, so it may be refilled by the next call)
BProb na_tab[NA_MAXPROB];
int na_ord[NA_MAXPROB];
std::pair<double, int> na_key[NA_MAXPROB];
}  // namespace

// once per device, with that device current: the array's device address
// and the persistent grid
static int na_setup(int dev) {
  if (na_ptr[dev]) return 0;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    (void)hipGetLastError();
    fprintf(stderr, "lila_hip: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
    return -1;
  }
  int ncu = 0, per = 0;
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)batched_ls_kernel, BT, 0);
  na_grid[dev] = (ncu > 0 ? ncu : 1) * (per > 0 ? per : 1);
  na_ptr[dev] = ptr;
  return 0;
}

extern "C" size_t ld_batch_noalloc_max(void) { return NA_MAXPROB; }

// set-up for the current device (the library's init call runs it)
extern "C" int ld_batch_noalloc_setup(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAXDEV) {
    (void)hipGetLastError();
    return -1;
  }
  return na_setup(dev);
}

// enqueue the batch on the library stream; 0, or -1 for arguments out of
// range (nothing launched)
extern "C" int ld_batch_noalloc_run(int nprob, const int *m, const int *n, const int *k,
                                    double *const *A, const int *lda, double *const *B,
                                    const int *ldb) {
  if (nprob == 0) return 0;
  if (nprob < 0 || nprob > NA_MAXPROB) {
    fprintf(stderr, "lila_hip: lila_dgels_batched_noalloc takes at most %d problems (got %d)\n",
            NA_MAXPROB, nprob);
    return -1;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAXDEV) {
    (void)hipGetLastError();
    return -1;
  }
  if (na_setup(dev) != 0) return -1;
  for (int i = 0; i < nprob; i++) {
    if (m[i] < n[i] || n[i] < 1 || n[i] > MAXN || k[i] < 0 || k[i] > MAXK || lda[i] < m[i] ||
        (k[i] > 0 && ldb[i] < m[i])) {
      fprintf(stderr, "lila_hip: batched problem %d out of range (m=%d n=%d k=%d)\n", i, m[i], n[i],
              k[i]);
      return -1;
    }
    na_tab[i] = BProb{m[i], n[i], k[i], lda[i], ldb[i], 0, A[i], B[i]};
    // queue order: largest flop count first, ties by index (the plan's order)
    na_key[i] = std::make_pair(-prob_flops(m[i], n[i], k[i]), i);
  }
  std::sort(na_key, na_key + nprob);  // in place, no allocation
  for (int i = 0; i < nprob; i++) na_ord[i] = na_key[i].second;
  hipStream_t st = (hipStream_t)ld_stream();
  char *base = (char *)na_ptr[dev];
  BProb *dprob = (BProb *)base;
  int *dorder = (int *)(base + NA_TAB);
  int *dcounter = (int *)(base + NA_TAB + NA_ORD);
  if (hipMemcpyAsync(dprob, na_tab, sizeof(BProb) * nprob, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(dorder, na_ord, sizeof(int) * nprob, hipMemcpyHostToDevice, st) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  const int grid = na_grid[dev] < nprob ? na_grid[dev] : nprob;
  // the counter is reset by a kernel, not a memset (the last call's value stays)
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(1), 0, st, dcounter);
  hipLaunchKernelGGL(batched_ls_kernel, dim3(grid), dim3(BT), 0, st, dprob, dorder, nprob, dcounter);
  if (hipGetLastError() != hipSuccess) return -1;
  return 0;
}

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
