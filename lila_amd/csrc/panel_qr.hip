// panel_qr.hip -- fused Householder panel for the LILA hot path (gfx950).
//
// One launch factors an ml x nb panel (nb <= 32): the row range is split
// into G slabs, one workgroup each, and every slab lives in LDS for the
// whole factorization.  Column j costs ONE grid-wide reduction: each slab
// contributes, in a single pass over its rows,
//   * the dlassq-style (scale, ssq) pair of x = A(j+1:, j),
//   * the dot products x^T A(j+1:, k) for k > j,
//   * the dot products v_p^T v_{j-1} (p < j-1) of the previous reflector
//     (the Gram entries that the T factor needs),
// and after the barrier every workgroup forms beta, tau and
//   v^T a_k = a_jk + (x^T a_k) / (alpha - beta)
// itself (identical arithmetic in every workgroup), so the reflector is
// generated (LAPACK dlarfg convention: beta = -sign(alpha) ||[alpha;x]||,
// tau = (beta - alpha)/beta, v = x/(alpha - beta)) and applied without a
// second barrier.  T = dlarft('F','C') is built from the Gram entries, and
// the explicit panel Q = [I;0] - V (T V1^T) is written in the same launch.
//
// Partial sums go through an agent-scope coherent buffer (atomic loads and
// stores) double-buffered by column parity; the reduction order is fixed,
// so results are deterministic.  The grid barrier has a bounded wait that
// raises a device flag instead of hanging.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include "lila_dev.h"

extern "C" int *ld_device_flags(void);

namespace {

constexpr int PNT = 256;
constexpr int PNB = LD_PANEL_MAX;  // 32
constexpr int SLOT = 128;
constexpr int RMAX = 480;          // rows per slab held in LDS

struct PanelArgs {
  int ml, nb;
  double *A; int lda;
  double *T; int ldt;
  double *Q; int ldq;
  int G, rp;
  double *part;       // 2 * G * SLOT partials + PNB*PNB V1 staging
  unsigned *cnt;      // [0] arrivals, [1] exits
  int *err;
};

__device__ __forceinline__ void st_coh(double *p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long *>(p),
                     (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_coh(const double *p) {
  unsigned long long u = __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)u);
}

__device__ void grid_barrier(unsigned *cnt, unsigned target, int *err) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(&cnt[0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (__hip_atomic_load(&cnt[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > (1u << 24)) {
        atomicOr(err, 1);
        break;
      }
    }
  }
  __syncthreads();
  __threadfence();
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// combine two (scale, ssq) pairs of the scaled sum of squares
__device__ __forceinline__ void ssq_combine(double &sc, double &sq, double sc2, double sq2) {
  if (sc2 == 0.0) return;
  if (sc == 0.0) { sc = sc2; sq = sq2; return; }
  if (sc >= sc2) { double r = sc2 / sc; sq += sq2 * r * r; }
  else { double r = sc / sc2; sq = sq * r * r + sq2; sc = sc2; }
}

__global__ __launch_bounds__(PNT) void panel_qr_kernel(PanelArgs a) {
  extern __shared__ double smem[];
  const int nb = a.nb, RP = a.rp, G = a.G;
  double *slab = smem;              // nb * RP, column-major
  double *Gm = slab + nb * RP;      // Gram, Gm[p + PNB*j]
  double *Tl = Gm + PNB * PNB;      // T
  double *Xl = Tl + PNB * PNB;      // V1, then T V1^T
  double *red = Xl + PNB * PNB;     // 4 * 64 scratch
  double *vals = red + 4 * 64;      // reduced p_k (0..63) and Gram (64..127)
  double *rowj = vals + 128;        // row j values, k = j..nb-1
  double *tauL = rowj + PNB;        // tau
  double *wk = tauL + PNB;          // w_k
  double *scal = wk + PNB;          // [0] beta [1] tau [2] scale [3] xnorm

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = blockIdx.x;
  const int r0 = g * RP;
  const int r1 = min(a.ml, r0 + RP);
  const int nloc = max(0, r1 - r0);

  for (int idx = tid; idx < nloc * nb; idx += PNT) {
    int lr = idx % nloc, c = idx / nloc;
    slab[lr + c * RP] = a.A[(size_t)(r0 + lr) + (size_t)c * a.lda];
  }
  __syncthreads();

  unsigned nbar = 0;
  for (int j = 0; j <= nb; j++) {
    const int nk = (j < nb) ? nb - 1 - j : 0;
    const int nr = (j < nb) ? nb - j : 0;
    const int ng = (j >= 1) ? j - 1 : 0;
    double *pp = a.part + ((size_t)(j & 1) * G + g) * SLOT;

    // ---- local partials: job 0 = norm (j < nb), then nk dots, then ng Gram
    const int njob = (j < nb ? 1 : 0) + nk + ng;
    for (int q = wave; q < njob; q += 4) {
      if (j < nb && q == 0) {
        double sc = 0.0, sq = 1.0;
        for (int lr = lane; lr < nloc; lr += 64) {
          if (r0 + lr <= j) continue;
          double x = fabs(slab[lr + j * RP]);
          if (x != 0.0) {
            if (sc < x) { double r = sc / x; sq = 1.0 + sq * r * r; sc = x; }
            else { double r = x / sc; sq += r * r; }
          }
        }
        if (sc == 0.0) sq = 0.0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          double sc2 = __shfl_xor(sc, o), sq2 = __shfl_xor(sq, o);
          ssq_combine(sc, sq, sc2, sq2);
        }
        if (lane == 0) { st_coh(&pp[0], sq); st_coh(&pp[1], sc); }
        continue;
      }
      int qq = q - (j < nb ? 1 : 0);
      double s = 0.0;
      if (qq < nk) {
        const int k = j + 1 + qq;
        for (int lr = lane; lr < nloc; lr += 64)
          if (r0 + lr > j) s += slab[lr + j * RP] * slab[lr + k * RP];
        s = wave_sum(s);
        if (lane == 0) st_coh(&pp[2 + qq], s);
      } else {
        const int p = qq - nk, c = j - 1;
        for (int lr = lane; lr < nloc; lr += 64) {
          int r = r0 + lr;
          if (r > c) s += slab[lr + p * RP] * slab[lr + c * RP];
          else if (r == c) s += slab[lr + p * RP];
        }
        s = wave_sum(s);
        if (lane == 0) st_coh(&pp[2 + nk + nr + (qq - nk)], s);
      }
    }
    if (g == 0) {
      for (int t = tid; t < nr; t += PNT) st_coh(&pp[2 + nk + t], slab[j + (j + t) * RP]);
    }
    if (j == nb && g == 0) {
      // stage V1 (strict lower of the top nb x nb) for every workgroup
      double *v1 = a.part + (size_t)2 * G * SLOT;
      for (int t = tid; t < nb * nb; t += PNT) {
        int r = t % nb, c = t / nb;
        st_coh(&v1[t], r > c ? slab[r + c * RP] : 0.0);
      }
    }
    nbar++;
    grid_barrier(a.cnt, nbar * (unsigned)G, a.err);

    // ---- reduce (same order in every workgroup) ----
    const double *pb = a.part + (size_t)(j & 1) * G * SLOT;
    const int nsum = nk + ng;
    {
      const int s = tid & 63, grp = tid >> 6;
      double acc = 0.0;
      if (s < nsum) {
        const int slot = (s < nk) ? 2 + s : 2 + nk + nr + (s - nk);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int w = grp;
        for (; w + 12 < G; w += 16) {
          a0 += ld_coh(&pb[(size_t)w * SLOT + slot]);
          a1 += ld_coh(&pb[(size_t)(w + 4) * SLOT + slot]);
          a2 += ld_coh(&pb[(size_t)(w + 8) * SLOT + slot]);
          a3 += ld_coh(&pb[(size_t)(w + 12) * SLOT + slot]);
        }
        for (; w < G; w += 4) a0 += ld_coh(&pb[(size_t)w * SLOT + slot]);
        acc = (a0 + a1) + (a2 + a3);
      }
      red[grp * 64 + s] = acc;
    }
    if (tid < nr) rowj[tid] = ld_coh(&pb[2 + nk + tid]);
    __syncthreads();
    if (tid < 64) {
      double v = red[tid] + red[64 + tid] + red[128 + tid] + red[192 + tid];
      if (tid < nk) vals[tid] = v;
      else if (tid < nsum) vals[64 + (tid - nk)] = v;
    }
    if (j < nb && wave == 1) {
      // norm: combine (scale, ssq) over workgroups in workgroup order
      double sc = 0.0, sq = 0.0;
      for (int w = lane; w < G; w += 64) {
        double sq2 = ld_coh(&pb[(size_t)w * SLOT + 0]);
        double sc2 = ld_coh(&pb[(size_t)w * SLOT + 1]);
        ssq_combine(sc, sq, sc2, sq2);
      }
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        double sc2 = __shfl_xor(sc, o), sq2 = __shfl_xor(sq, o);
        ssq_combine(sc, sq, sc2, sq2);
      }
      if (lane == 0) scal[3] = (sc == 0.0) ? 0.0 : sc * sqrt(sq);
    }
    __syncthreads();
    if (tid == 0 && j < nb) {
      const double alpha = rowj[0], xnorm = scal[3];
      double beta = alpha, tau = 0.0, scale = 0.0;
      if (xnorm != 0.0) {
        double aa = fabs(alpha);
        double w = fmax(aa, xnorm), z = fmin(aa, xnorm);
        double nrm = (z == 0.0) ? w : w * sqrt(1.0 + (z / w) * (z / w));
        beta = -copysign(nrm, alpha);
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      scal[0] = beta; scal[1] = tau; scal[2] = scale;
      tauL[j] = tau;
    }
    __syncthreads();
    if (j < nb && tid > j && tid < nb) {
      const int k = tid;
      wk[k] = scal[1] * (rowj[k - j] + vals[k - j - 1] * scal[2]);
    }
    if (j >= 1 && tid < j - 1) Gm[tid + PNB * (j - 1)] = vals[64 + tid];
    __syncthreads();

    // ---- apply the reflector to the slab ----
    if (j < nb && scal[1] != 0.0) {
      const double scale = scal[2];
      for (int lr = tid; lr < nloc; lr += PNT) {
        const int r = r0 + lr;
        if (r > j) {
          double v = slab[lr + j * RP] * scale;
          slab[lr + j * RP] = v;
          for (int k = j + 1; k < nb; k++) slab[lr + k * RP] -= v * wk[k];
        } else if (r == j) {
          slab[lr + j * RP] = scal[0];
          for (int k = j + 1; k < nb; k++) slab[lr + k * RP] -= wk[k];
        }
      }
    }
    __syncthreads();
  }

  // ---- write V and R back ----
  for (int idx = tid; idx < nloc * nb; idx += PNT) {
    int lr = idx % nloc, c = idx / nloc;
    a.A[(size_t)(r0 + lr) + (size_t)c * a.lda] = slab[lr + c * RP];
  }

  // ---- T = dlarft('F','C') from tau and the Gram entries ----
  for (int t = tid; t < PNB * PNB; t += PNT) Tl[t] = 0.0;
  __syncthreads();
  for (int j = 0; j < nb; j++) {
    const double tj = tauL[j];
    if (tid < j) red[tid] = -tj * Gm[tid + PNB * j];
    __syncthreads();
    if (tid < j) {
      double s = 0.0;
      if (tj != 0.0)
        for (int q = tid; q < j; q++) s += Tl[tid + PNB * q] * red[q];
      Tl[tid + PNB * j] = s;
    }
    if (tid == j) Tl[j + PNB * j] = tj;
    __syncthreads();
  }
  if (g == 0) {
    for (int t = tid; t < nb * nb; t += PNT) {
      int r = t % nb, c = t / nb;
      if (r <= c) a.T[(size_t)r + (size_t)c * a.ldt] = Tl[r + PNB * c];
    }
  }

  // ---- explicit panel Q = [I;0] - V (T V1^T) ----
  if (a.Q) {
    const double *v1 = a.part + (size_t)2 * G * SLOT;
    for (int t = tid; t < nb * nb; t += PNT) {
      int r = t % nb, c = t / nb;
      Xl[r + PNB * c] = ld_coh(&v1[t]);  // V1 strict lower
    }
    __syncthreads();
    // X(r,c) = sum_{p=r..c} T(r,p) V1(c,p), V1(c,c) = 1; stored over V1's
    // upper part (r <= c), which V1 does not use
    double xv[4];
    int cntx = 0;
    for (int t = tid; t < nb * nb; t += PNT) {
      int r = t % nb, c = t / nb;
      double s = 0.0;
      if (r <= c) {
        s = Tl[r + PNB * c];
        for (int p = r; p < c; p++) s += Tl[r + PNB * p] * Xl[c + PNB * p];
      }
      xv[cntx++] = s;
    }
    __syncthreads();
    cntx = 0;
    for (int t = tid; t < nb * nb; t += PNT) {
      int r = t % nb, c = t / nb;
      if (r <= c) Xl[r + PNB * c] = xv[cntx];
      cntx++;
    }
    __syncthreads();
    for (int idx = tid; idx < nloc * nb; idx += PNT) {
      const int lr = idx % nloc, c = idx / nloc, r = r0 + lr;
      double s = (r == c) ? 1.0 : 0.0;
      const int pmax = min(r, c);
      for (int p = 0; p <= pmax; p++) {
        double v = (r == p) ? 1.0 : slab[lr + p * RP];
        s -= v * Xl[p + PNB * c];
      }
      a.Q[(size_t)r + (size_t)c * a.ldq] = s;
    }
  }

  // ---- last workgroup out resets the barrier counters ----
  __syncthreads();
  if (tid == 0) {
    unsigned old = __hip_atomic_fetch_add(&a.cnt[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old == (unsigned)G - 1) {
      __hip_atomic_store(&a.cnt[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&a.cnt[1], 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

size_t panel_smem_bytes(int nb, int rp) {
  return sizeof(double) * ((size_t)nb * rp + 3 * PNB * PNB + 4 * 64 + 128 + 4 * PNB + 8);
}

}  // namespace

extern "C" void ld_panel_qr(int ml, int nb, double *A, int lda, double *T, int ldt,
                            double *Q, int ldq) {
  if (nb <= 0 || ml <= 0) return;
  if (nb > PNB) {
    ld_fatal("panel width above 32 reached the fused panel kernel");
    return;
  }
  int target = ld_tune()->panel_rows_per_wg;
  if (target < nb) target = nb;
  if (target > RMAX) target = RMAX;
  int G = (ml + target - 1) / target;
  if (G > 256) G = 256;
  int rp = (ml + G - 1) / G;
  if (rp < nb) rp = nb;
  G = (ml + rp - 1) / rp;
  if (rp > RMAX) {
    ld_fatal("panel taller than 256 x 512 rows is not supported by the fused panel kernel");
    return;
  }
  static int attr_set = 0;
  if (!attr_set) {
    hipFuncSetAttribute((const void *)panel_qr_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)panel_smem_bytes(PNB, RMAX));
    attr_set = 1;
  }
  size_t mark = ld_ws_mark();
  PanelArgs pa;
  pa.ml = ml; pa.nb = nb; pa.A = A; pa.lda = lda; pa.T = T; pa.ldt = ldt;
  pa.Q = Q; pa.ldq = ldq; pa.G = G; pa.rp = rp;
  pa.part = (double *)ld_ws_alloc(sizeof(double) * ((size_t)2 * G * SLOT + PNB * PNB));
  int *flags = ld_device_flags();
  pa.cnt = (unsigned *)(flags + 8);
  pa.err = flags;
  void *args[] = {&pa};
  const int ph = ld_prof_begin("panel_qr_kernel");
  hipError_t e = hipLaunchCooperativeKernel((const void *)panel_qr_kernel, dim3(G), dim3(PNT), args,
                                            (unsigned)panel_smem_bytes(nb, rp),
                                            (hipStream_t)ld_stream());
  // algorithmic traffic: read the panel, write V/R (+ Q), write T
  ld_prof_end(ph, 2.0 * ml * nb * nb + (Q ? 2.0 * ml * nb * nb : 0.0),
              8.0 * ((Q ? 3.0 : 2.0) * ml * nb + (double)nb * nb));
  if (e != hipSuccess) {
    fprintf(stderr, "lila_hip: panel launch failed (G=%d rp=%d nb=%d): %s\n", G, rp, nb,
            hipGetErrorString(e));
    ld_set_info(-1);
  }
  ld_ws_release(mark);
}
