// panel_qr.hip -- fused Householder panel for the LILA hot path (gfx950).
//
// One launch factors an ml x nb panel (nb <= 32): the row range is split
// into G slabs, one workgroup each, and every slab lives in LDS for the
// whole factorization.  Column j costs ONE grid-wide reduction: each slab
// contributes, in a single pass over its rows,
//   * sum and max of x^2 for x = A(j+1:, j)  (the norm; the max guards the
//     plain sum of squares against over/underflow, with a scaled second
//     pass when it is outside [1e-150, 1e150]),
//   * the dot products x^T A(j+1:, k) for k > j,
//   * the dot products v_p^T v_{j-1} (p < j-1) of the previous reflector
//     (the Gram entries the T factor needs),
// accumulated with hardware FP64 atomic adds into a triple-buffered vector,
// and after the barrier every workgroup forms beta, tau and
//   v^T a_k = a_jk + (x^T a_k) / (alpha - beta)
// itself (identical arithmetic in every workgroup), so the reflector is
// generated (LAPACK dlarfg convention: beta = -sign(alpha) ||[alpha;x]||,
// tau = (beta - alpha)/beta, v = x/(alpha - beta)) and applied without a
// second barrier.  T = dlarft('F','C') is built from the Gram entries, and
// the explicit panel Q = [I;0] - V (T V1^T) is written in the same launch.
//
// Cross-workgroup data only moves through agent-scope atomics (adds, max,
// coherent loads/stores); a barrier is one relaxed add and a poll by one
// thread per workgroup after an s_waitcnt, with a bounded wait that raises
// a device flag instead of hanging.  The atomic sums make the reduction order (and
// so the last bits of the result) vary from run to run.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include "lila_dev.h"

extern "C" int *ld_device_flags(void);
extern "C" double *ld_panel_scratch(void);

namespace {

constexpr int PNT = 256;
constexpr int PNB = LD_PANEL_MAX;  // 32
constexpr int SLOT = 128;          // one reduction vector
constexpr int SLOT_SCALED = 126;   // slot of the scaled sum of squares
constexpr int RMAX = 480;          // rows per slab held in LDS

struct PanelArgs {
  int ml, nb;
  double *A; int lda;
  double *T; int ldt;
  double *Q; int ldq;
  int G, rp;
  double *red;        // 3 * SLOT reduction vectors + PNB*PNB V1 staging
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
__device__ __forceinline__ void add_coh(double *p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void max_coh(double *p, double v) {  // v >= 0
  __hip_atomic_fetch_max(reinterpret_cast<unsigned long long *>(p),
                         (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// All waves' atomics complete (s_waitcnt), then one arrival per workgroup.
// Every value exchanged between workgroups is itself an agent-scope atomic
// performed at the coherence point, so no cache writeback / invalidate
// fences are needed: the arrival is issued only after this workgroup's
// atomics completed, and the reads of the reduced values are issued only
// after the poll saw every arrival.
__device__ void grid_sync(unsigned *cnt, unsigned target, int *err) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(&cnt[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (__hip_atomic_load(&cnt[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > (1u << 24)) {
        atomicOr(err, 1);
        break;
      }
    }
  }
  __syncthreads();
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

__global__ __launch_bounds__(PNT) void panel_qr_kernel(PanelArgs a) {
  extern __shared__ double smem[];
  const int nb = a.nb, RP = a.rp, G = a.G;
  double *slab = smem;              // nb * RP, column-major
  double *Gm = slab + nb * RP;      // Gram, Gm[p + PNB*j]
  double *Tl = Gm + PNB * PNB;      // T
  double *Xl = Tl + PNB * PNB;      // V1, then T V1^T
  double *tmp = Xl + PNB * PNB;     // 64 scratch
  double *vals = tmp + 64;          // reduced p_k (0..63) and Gram (64..127)
  double *rowj = vals + 128;        // row j values, k = j..nb-1
  double *tauL = rowj + PNB;        // tau
  double *wk = tauL + PNB;          // w_k
  double *scal = wk + PNB;          // [0] beta [1] tau [2] scale [3] xnorm [4] maxabs

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
    double *rv = a.red + (size_t)(j % 3) * SLOT;
    if (g == 0 && tid < SLOT) st_coh(&a.red[(size_t)((j + 1) % 3) * SLOT + tid], 0.0);

    // ---- local sums: job 0 = norm (j < nb), then nk dots, then ng Gram
    const int njob = (j < nb ? 1 : 0) + nk + ng;
    for (int q = wave; q < njob; q += 4) {
      if (j < nb && q == 0) {
        double s = 0.0, mx = 0.0;
        for (int lr = lane; lr < nloc; lr += 64) {
          if (r0 + lr <= j) continue;
          double x = slab[lr + j * RP];
          s += x * x;
          mx = fmax(mx, fabs(x));
        }
        s = wave_sum(s);
        mx = wave_max(mx);
        if (lane == 0) {
          if (s != 0.0) add_coh(&rv[0], s);
          if (mx != 0.0) max_coh(&rv[1], mx);
        }
        continue;
      }
      const int qq = q - (j < nb ? 1 : 0);
      double s = 0.0;
      if (qq < nk) {
        const int k = j + 1 + qq;
        for (int lr = lane; lr < nloc; lr += 64)
          if (r0 + lr > j) s += slab[lr + j * RP] * slab[lr + k * RP];
        s = wave_sum(s);
        if (lane == 0 && s != 0.0) add_coh(&rv[2 + qq], s);
      } else {
        const int p = qq - nk, c = j - 1;
        for (int lr = lane; lr < nloc; lr += 64) {
          const int r = r0 + lr;
          if (r > c) s += slab[lr + p * RP] * slab[lr + c * RP];
          else if (r == c) s += slab[lr + p * RP];
        }
        s = wave_sum(s);
        if (lane == 0 && s != 0.0) add_coh(&rv[2 + nk + nr + p], s);
      }
    }
    if (g == 0) {
      for (int t = tid; t < nr; t += PNT) st_coh(&rv[2 + nk + t], slab[j + (j + t) * RP]);
    }
    if (j == nb && g == 0) {
      // stage V1 (strict lower of the top nb x nb) for every workgroup
      double *v1 = a.red + 3 * SLOT;
      for (int t = tid; t < nb * nb; t += PNT) {
        const int r = t % nb, c = t / nb;
        st_coh(&v1[t], r > c ? slab[r + c * RP] : 0.0);
      }
    }
    nbar++;
    grid_sync(a.cnt, nbar * (unsigned)G, a.err);

    // ---- read the reduced vector ----
    const int nsum = nk + ng;
    if (tid < nsum) {
      const int slot = (tid < nk) ? 2 + tid : 2 + nk + nr + (tid - nk);
      const double v = ld_coh(&rv[slot]);
      if (tid < nk) vals[tid] = v;
      else vals[64 + (tid - nk)] = v;
    }
    if (tid >= 64 && tid < 64 + nr) rowj[tid - 64] = ld_coh(&rv[2 + nk + (tid - 64)]);
    if (tid == 128 && j < nb) {
      scal[3] = ld_coh(&rv[0]);
      scal[4] = ld_coh(&rv[1]);
    }
    __syncthreads();
    if (j < nb) {
      const double mx = scal[4];
      if (mx != 0.0 && (mx < 1e-150 || mx > 1e150)) {
        // scaled second pass: sum of (x / max)^2
        if (wave == 0) {
          const double inv = 1.0 / mx;
          double s = 0.0;
          for (int lr = lane; lr < nloc; lr += 64) {
            if (r0 + lr <= j) continue;
            const double x = slab[lr + j * RP] * inv;
            s += x * x;
          }
          s = wave_sum(s);
          if (lane == 0 && s != 0.0) add_coh(&rv[SLOT_SCALED], s);
        }
        nbar++;
        grid_sync(a.cnt, nbar * (unsigned)G, a.err);
        if (tid == 0) scal[3] = mx * sqrt(ld_coh(&rv[SLOT_SCALED]));
      } else if (tid == 0) {
        scal[3] = sqrt(scal[3]);
      }
      __syncthreads();
    }
    if (tid == 0 && j < nb) {
      const double alpha = rowj[0], xnorm = scal[3];
      double beta = alpha, tau = 0.0, scale = 0.0;
      if (xnorm != 0.0) {
        const double aa = fabs(alpha);
        const double w = fmax(aa, xnorm), z = fmin(aa, xnorm);
        const double nrm = (z == 0.0) ? w : w * sqrt(1.0 + (z / w) * (z / w));
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
          const double v = slab[lr + j * RP] * scale;
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
    const int lr = idx % nloc, c = idx / nloc;
    a.A[(size_t)(r0 + lr) + (size_t)c * a.lda] = slab[lr + c * RP];
  }

  // ---- T = dlarft('F','C') from tau and the Gram entries ----
  for (int t = tid; t < PNB * PNB; t += PNT) Tl[t] = 0.0;
  __syncthreads();
  for (int j = 0; j < nb; j++) {
    const double tj = tauL[j];
    if (tid < j) tmp[tid] = -tj * Gm[tid + PNB * j];
    __syncthreads();
    if (tid < j) {
      double s = 0.0;
      if (tj != 0.0)
        for (int q = tid; q < j; q++) s += Tl[tid + PNB * q] * tmp[q];
      Tl[tid + PNB * j] = s;
    }
    if (tid == j) Tl[j + PNB * j] = tj;
    __syncthreads();
  }
  if (g == 0) {
    for (int t = tid; t < nb * nb; t += PNT) {
      const int r = t % nb, c = t / nb;
      if (r <= c) a.T[(size_t)r + (size_t)c * a.ldt] = Tl[r + PNB * c];
    }
  }

  // ---- explicit panel Q = [I;0] - V (T V1^T) ----
  if (a.Q) {
    const double *v1 = a.red + 3 * SLOT;
    for (int t = tid; t < nb * nb; t += PNT) {
      const int r = t % nb, c = t / nb;
      Xl[r + PNB * c] = ld_coh(&v1[t]);  // V1 strict lower
    }
    __syncthreads();
    // X(r,c) = sum_{p=r..c} T(r,p) V1(c,p), V1(c,c) = 1; stored over V1's
    // upper part (r <= c), which V1 does not use
    double xv[4];
    int cntx = 0;
    for (int t = tid; t < nb * nb; t += PNT) {
      const int r = t % nb, c = t / nb;
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
      const int r = t % nb, c = t / nb;
      if (r <= c) Xl[r + PNB * c] = xv[cntx];
      cntx++;
    }
    __syncthreads();
    for (int idx = tid; idx < nloc * nb; idx += PNT) {
      const int lr = idx % nloc, c = idx / nloc, r = r0 + lr;
      double s = (r == c) ? 1.0 : 0.0;
      const int pmax = min(r, c);
      for (int p = 0; p <= pmax; p++) {
        const double v = (r == p) ? 1.0 : slab[lr + p * RP];
        s -= v * Xl[p + PNB * c];
      }
      a.Q[(size_t)r + (size_t)c * a.ldq] = s;
    }
  }

  // ---- last workgroup out re-zeroes the reduction vectors and counters ----
  __syncthreads();
  __shared__ unsigned last;
  if (tid == 0) {
    const unsigned old = __hip_atomic_fetch_add(&a.cnt[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (old == (unsigned)G - 1) ? 1u : 0u;
  }
  __syncthreads();
  if (last) {
    for (int t = tid; t < 3 * SLOT; t += PNT) st_coh(&a.red[t], 0.0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) {
      __hip_atomic_store(&a.cnt[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&a.cnt[1], 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

size_t panel_smem_bytes(int nb, int rp) {
  return sizeof(double) * ((size_t)nb * rp + 3 * PNB * PNB + 64 + 128 + 3 * PNB + 8);
}

}  // namespace

// persistent scratch size (doubles) the runtime allocates and zeroes once
extern "C" size_t ld_panel_scratch_doubles(void) { return 3 * SLOT + PNB * PNB; }

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
    ld_fatal("panel taller than 256 x 480 rows is not supported by the fused panel kernel");
    return;
  }
  // Co-residency of all G workgroups is what the in-kernel barrier needs.
  // It is checked once from the occupancy calculator (one workgroup per CU
  // at the largest LDS size, G <= CU count); a plain launch then keeps all
  // G resident because the stream runs nothing else meanwhile.  The
  // cooperative launch API gives the same guarantee at a higher launch cost
  // and is used when LILA_PANEL_COOP=1 or when the check fails.
  static int attr_set = 0, ncu = 0, coop = 0;
  if (!attr_set) {
    hipFuncSetAttribute((const void *)panel_qr_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)panel_smem_bytes(PNB, RMAX));
    int dev = 0, per_cu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)panel_qr_kernel, PNT,
                                                 panel_smem_bytes(PNB, RMAX));
    const char *e = getenv("LILA_PANEL_COOP");
    coop = (e && atoi(e) == 1) || per_cu < 1 || ncu < 1;
    attr_set = 1;
  }
  if (ncu > 0 && G > ncu) {
    ld_fatal("panel needs more workgroups than compute units");
    return;
  }
  PanelArgs pa;
  pa.ml = ml; pa.nb = nb; pa.A = A; pa.lda = lda; pa.T = T; pa.ldt = ldt;
  pa.Q = Q; pa.ldq = ldq; pa.G = G; pa.rp = rp;
  pa.red = ld_panel_scratch();
  int *flags = ld_device_flags();
  pa.cnt = (unsigned *)(flags + 8);
  pa.err = flags;
  void *args[] = {&pa};
  const int ph = ld_prof_begin("panel_qr_kernel");
  hipError_t e;
  if (coop) {
    e = hipLaunchCooperativeKernel((const void *)panel_qr_kernel, dim3(G), dim3(PNT), args,
                                   (unsigned)panel_smem_bytes(nb, rp), (hipStream_t)ld_stream());
  } else {
    hipLaunchKernelGGL(panel_qr_kernel, dim3(G), dim3(PNT), panel_smem_bytes(nb, rp),
                       (hipStream_t)ld_stream(), pa);
    e = hipGetLastError();
  }
  // algorithmic traffic: read the panel, write V/R (+ Q), write T
  ld_prof_end(ph, 2.0 * ml * nb * nb + (Q ? 2.0 * ml * nb * nb : 0.0),
              8.0 * ((Q ? 3.0 : 2.0) * ml * nb + (double)nb * nb));
  if (e != hipSuccess) {
    fprintf(stderr, "lila_hip: panel launch failed (G=%d rp=%d nb=%d): %s\n", G, rp, nb,
            hipGetErrorString(e));
    ld_set_info(-1);
  }
}
