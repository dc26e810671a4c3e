// panel_qr.hip -- fused Householder panels for the LILA hot path (gfx950).
//
// An ml x nb panel (nb <= 32) is split into G row slabs, one workgroup
// each; every slab lives in LDS for the whole launch.  Two kernels:
//
// 1. panel_hh_kernel -- column-by-column Householder (dgeqr3's numbers):
//    column j costs ONE grid-wide reduction: each slab contributes, in one
//    pass over its rows, sum and max of x^2 for x = A(j+1:, j) (the max
//    guards the plain sum of squares, with a scaled second pass outside
//    [1e-150, 1e150]), the dot products x^T A(j+1:, k) for k > j, and the
//    Gram entries v_p^T v_{j-1} the T factor needs; every workgroup then
//    forms beta, tau (LAPACK dlarfg convention: beta = -sign(alpha)
//    ||[alpha;x]||, tau = (beta-alpha)/beta, v = x/(alpha-beta)) and
//    v^T a_k = a_jk + (x^T a_k)/(alpha - beta) itself and applies the
//    reflector.  T = dlarft('F','C') comes from the Gram entries and the
//    explicit panel Q = [I;0] - V (T V1^T) is written in the same launch.
//    With G = 1 the reductions stay in LDS (no global traffic).
//
// 2. panel_cqr_kernel -- communication-avoiding: Cholesky-QR twice
//    (Gram X^T X of each slab on v_mfma_f64_16x16x4, one grid reduction per
//    pass, the 32 x 32 Cholesky in registers of one wave, Q = X R^-1 as an
//    MFMA product with the explicit inverse), then Householder
//    reconstruction: sign-chosen LU of the top nb x nb block of Q
//    (Q1 S - I = L U, formed by every workgroup from its own copy of the
//    top rows), V1 = L, V2 = Q2 S U^-1, T = -U V1^-T, R = S R2 R1,
//    explicit Q = Q S.  Two grid reductions per panel instead of nb + 1.  In exact arithmetic this is the same V, R, T
//    and Q as the column-by-column Householder panel (Q and R are unique up
//    to column signs and the reconstruction picks dlarfg's signs).  If a
//    Cholesky pivot is not safely positive, the diagonal of R spreads by
//    more than 1e5, or a reflector has tau within 1e-10 of 0 or 2, every
//    workgroup reloads its slab and runs the column-by-column loop instead.
//
// Cross-workgroup data only moves through agent-scope atomics (FP64 adds,
// max, coherent loads/stores); a barrier is an s_waitcnt, one relaxed add
// and a poll by one thread per workgroup, with a bounded wait that raises
// a device flag instead of hanging.  The atomic sums make the reduction
// order (and so the last bits of the result) vary from run to run.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "lila_dev.h"
#include "small32.h"
#include "gridsync.h"

extern "C" int *ld_device_flags(void);
extern "C" double *ld_panel_scratch(void);

typedef double d4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int PNT = 256;
constexpr int PNB = LD_PANEL_MAX;  // 32
constexpr int SLOT = 128;          // one reduction vector
constexpr int SLOT_SCALED = 126;   // slot of the scaled sum of squares
constexpr int SLOT_SDOT = 66;      // scaled dot products (a column's sums and row take slots 0..64)
static_assert(SLOT_SDOT > 2 * PNB && SLOT_SDOT + PNB - 1 < SLOT_SCALED, "reduction slot layout");
constexpr int LDM = 34;            // leading dimension of 32 x 32 LDS matrices (even: 16-byte pairs)
static_assert(PNB == small32::PNB && LDM == small32::LDM, "32 x 32 helpers");
using small32::rdlane;
using small32::wave_chol32;
using small32::wave_lu_sign32;
using small32::wfence;
using small32::wg_mm32;
// persistent global scratch layout (doubles)
constexpr int OFF_RED = 0;                   // 3 x SLOT reduction vectors
constexpr int OFF_V1 = OFF_RED + 3 * SLOT;   // V1 staging (PNB*PNB)
constexpr int OFF_G1 = OFF_V1 + PNB * PNB;   // Gram, pass 1 (PNB*PNB)
constexpr int OFF_G2 = OFF_G1 + PNB * PNB;   // Gram, pass 2
constexpr int OFF_PUB = OFF_G2 + PNB * PNB;  // U^-1 (PNB*PNB), S (PNB), flag
constexpr int SCRATCH = OFF_PUB + PNB * PNB + PNB + 8;
// LDS past the slab (doubles): the larger of the two kernels' needs
constexpr int HH_TAIL = 3 * PNB * PNB + 64 + 128 + 3 * PNB + 16 + SLOT;
constexpr int CQR_TAIL = 4 * PNB * LDM + 3 * PNB + 2;
constexpr int LDS_TAIL = (HH_TAIL > CQR_TAIL ? HH_TAIL : CQR_TAIL) + 2;
// streamed panel (taller than the workgroups' LDS): five 32 x 32 matrices
constexpr int STREAM_TAIL = 5 * PNB * LDM + 3 * PNB + 4;
// dynamic LDS budget: 160 KiB minus the kernels' static LDS
constexpr int LDS_DYN_BYTES = 163840 - 16;
constexpr int LDS_SLAB_DOUBLES = LDS_DYN_BYTES / 8 - LDS_TAIL;

struct PanelArgs {
  int ml, nb;
  double *A; int lda;
  double *T; int ldt;
  double *Q; int ldq;
  int G, rp, rs;      // workgroups, rows per slab, LDS column stride
  int sr;             // streamed panel: rows per LDS sub-slab
  double *scr;        // persistent global scratch (SCRATCH doubles)
  unsigned *cnt;      // [1] exits
  unsigned *bar;      // grid barrier counters (in the persistent scratch)
  double *scr_next;   // 64-wide panel: the other parity's scratch set (zeroed here for the next launch)
  int *err;           // [0] barrier time-out, [2] a streamed panel declined
  int decline;        // streamed panel: leave a panel that fails the acceptance
                      // tests untouched and raise err[2] (no column loop)
  int batch;          // every workgroup factors a slab of its own (TSQR leaves)
};

using namespace gridsync;


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

// slab size in doubles, rounded up to keep what follows 16-byte aligned
__device__ __forceinline__ size_t slab_doubles(int nb, int rs) { return ((size_t)nb * rs + 1) & ~(size_t)1; }

// rows [r0, r0 + nloc) of an nb-column block of a column-major matrix into
// the slab and back: one row per thread, eight columns in flight
// the slab's nloc x nb entries, every thread taking entries nloc apart
// (loads coalesced over the rows, all of a thread's loads independent)
__device__ void load_slab(const PanelArgs &a, double *slab, int r0, int nloc) {
  // a wave's lanes run down the rows of a column (coalesced); each round
  // issues the loads of 8 columns x 2 blocks of 64 rows before its first
  // LDS write, so the round trips to HBM overlap
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nb = a.nb;
  for (int rb0 = 0; rb0 < nloc; rb0 += 128) {
    for (int c0 = wave; c0 < nb; c0 += 32) {
      double v[8][2];
#pragma unroll
      for (int q = 0; q < 8; q++)
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int c = c0 + 4 * q, lr = rb0 + 64 * h + lane;
          v[q][h] = (c < nb && lr < nloc) ? a.A[(size_t)(r0 + lr) + (size_t)c * a.lda] : 0.0;
        }
#pragma unroll
      for (int q = 0; q < 8; q++)
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int c = c0 + 4 * q, lr = rb0 + 64 * h + lane;
          if (c < nb && lr < nloc) slab[lr + c * a.rs] = v[q][h];
        }
    }
  }
  __syncthreads();
}
__device__ void store_slab(const double *slab, int rs, int nb, int nloc, double *dst, int ld,
                           const double *colscale) {
  for (int lr = threadIdx.x; lr < nloc; lr += PNT) {
    double *d = dst + lr;
#pragma unroll 8
    for (int c = 0; c < nb; c++)
      d[(size_t)c * ld] = colscale ? slab[lr + c * rs] * colscale[c] : slab[lr + c * rs];
  }
}

// ---------------------------------------------------------------------------
// column-by-column Householder on the slab (+ T, Q, write-back).
// SOLO: G == 1, reductions stay in LDS.  GLOB: the slab is the panel itself
// in global memory (rows r0.. of A, column stride lda): panels taller than
// the workgroups' LDS can hold stream every column step through HBM / L2.
// ---------------------------------------------------------------------------
template <bool SOLO, bool GLOB = false>
__device__ void hh_factor(const PanelArgs &a, double *slab, double *tail, int r0, int nloc,
                          unsigned &nbar) {
  const int nb = a.nb, RS = a.rs, G = a.G;
  double *Gm = tail;                // Gram, Gm[p + PNB*j]
  double *Tl = Gm + PNB * PNB;      // T
  double *Xl = Tl + PNB * PNB;      // V1, then T V1^T
  double *tmp = Xl + PNB * PNB;     // 64 scratch
  double *vals = tmp + 64;          // reduced p_k (0..63) and Gram (64..127)
  double *rowj = vals + 128;        // row j values, k = j..nb-1
  double *tauL = rowj + PNB;        // tau
  double *wk = tauL + PNB;          // w_k
  double *scal = wk + PNB;          // [0] beta [1] tau [2] scale [3] xnorm [4] maxabs
                                    // [5] scale of the dots [6] their factor in w_k
                                    // [7] a dot is not finite
  double *lred = scal + 16;         // SOLO: local reduction vector (SLOT)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = a.batch ? 0 : blockIdx.x;
  double *gred = a.scr + OFF_RED;
  if (tid == 0) scal[7] = 0.0;

  for (int j = 0; j <= nb; j++) {
    const int nk = (j < nb) ? nb - 1 - j : 0;
    const int nr = (j < nb) ? nb - j : 0;
    const int ng = (j >= 1) ? j - 1 : 0;
    double *rv = SOLO ? lred : gred + (size_t)(j % 3) * SLOT;
    if (!SOLO && g == 0 && tid < SLOT) st_coh(&gred[(size_t)((j + 1) % 3) * SLOT + tid], 0.0);

    // ---- local sums: job 0 = norm (j < nb), then nk dots, then ng Gram
    const int njob = (j < nb ? 1 : 0) + nk + ng;
    for (int q = wave; q < njob; q += 4) {
      if (j < nb && q == 0) {
        double s = 0.0, mx = 0.0;
        for (int lr = lane; lr < nloc; lr += 64) {
          if (r0 + lr <= j) continue;
          const double x = slab[lr + j * RS];
          s += x * x;
          mx = fmax(mx, fabs(x));
        }
        s = wave_sum(s);
        mx = wave_max(mx);
        if (lane == 0) {
          if (SOLO) {
            rv[0] = s;
            rv[1] = mx;
          } else {
            if (s != 0.0) add_coh(&rv[0], s);
            if (mx != 0.0) max_coh(&rv[1], mx);
          }
        }
        continue;
      }
      const int qq = q - (j < nb ? 1 : 0);
      double s = 0.0;
      int slot;
      if (qq < nk) {
        const int k = j + 1 + qq;
        for (int lr = lane; lr < nloc; lr += 64)
          if (r0 + lr > j) s += slab[lr + j * RS] * slab[lr + k * RS];
        slot = 2 + qq;
      } else {
        const int p = qq - nk, c = j - 1;
        for (int lr = lane; lr < nloc; lr += 64) {
          const int r = r0 + lr;
          if (r > c) s += slab[lr + p * RS] * slab[lr + c * RS];
          else if (r == c) s += slab[lr + p * RS];
        }
        slot = 2 + nk + nr + p;
      }
      s = wave_sum(s);
      if (lane == 0) {
        if (SOLO) rv[slot] = s;
        else if (s != 0.0) add_coh(&rv[slot], s);
      }
    }
    if (g == 0) {
      for (int t = tid; t < nr; t += PNT) {
        if (SOLO) rv[2 + nk + t] = slab[j + (j + t) * RS];
        else st_coh(&rv[2 + nk + t], slab[j + (j + t) * RS]);
      }
    }
    if (j == nb && g == 0 && !SOLO) {
      // stage V1 (strict lower of the top nb x nb) for every workgroup
      double *v1 = a.scr + OFF_V1;
      for (int t = tid; t < nb * nb; t += PNT) {
        const int r = t % nb, c = t / nb;
        st_coh(&v1[t], r > c ? slab[r + c * RS] : 0.0);
      }
    }
    if (SOLO) {
      __syncthreads();
    } else {
      nbar++;
      grid_sync(a.bar, nbar * (unsigned)G, a.err);
    }

    // ---- read the reduced vector ----
    const int nsum = nk + ng;
    if (tid < nsum) {
      const int slot = (tid < nk) ? 2 + tid : 2 + nk + nr + (tid - nk);
      const double v = SOLO ? rv[slot] : ld_coh(&rv[slot]);
      if (tid < nk) {
        vals[tid] = v;
        if (!(fabs(v) <= 1.7976931348623157e308)) scal[7] = 1.0;
      } else {
        vals[64 + (tid - nk)] = v;
      }
    }
    if (tid >= 64 && tid < 64 + nr)
      rowj[tid - 64] = SOLO ? rv[2 + nk + (tid - 64)] : ld_coh(&rv[2 + nk + (tid - 64)]);
    if (tid == 128 && j < nb) {
      scal[3] = SOLO ? rv[0] : ld_coh(&rv[0]);
      scal[4] = SOLO ? rv[1] : ld_coh(&rv[1]);
    }
    __syncthreads();
    if (j < nb) {
      const double mx = scal[4];
      // a product x_r a_rk out of range leaves a dot non-finite (or, with
      // the column's entries tiny, flushes it toward zero: mx < 1e-150)
      // (scal[7]: set while the sums were read, cleared again before the
      // next column's; no __syncthreads_or, whose LDS word would not fit
      // beside the dynamic LDS the launch asks for)
      const bool anybad = scal[7] != 0.0;
      if (mx != 0.0 && (mx < 1e-150 || mx > 1e150 || anybad)) {
        // scaled second pass (the purpose of dnrm2's scaling): the column
        // times a power of two near 1 / max|x|, exact; its sum of squares
        // and its dot products with the later columns
        const double inv = ldexp(1.0, -ilogb(mx)), sc = ldexp(1.0, ilogb(mx));
        for (int q = wave; q < 1 + nk; q += 4) {
          double sum = 0.0;
          if (q == 0) {
            for (int lr = lane; lr < nloc; lr += 64) {
              if (r0 + lr <= j) continue;
              const double x = slab[lr + j * RS] * inv;
              sum += x * x;
            }
          } else {
            const int k = j + q;
            for (int lr = lane; lr < nloc; lr += 64)
              if (r0 + lr > j) sum += (slab[lr + j * RS] * inv) * slab[lr + k * RS];
          }
          sum = wave_sum(sum);
          if (lane == 0) {
            const int slot = (q == 0) ? SLOT_SCALED : SLOT_SDOT + q - 1;
            if (SOLO) rv[slot] = sum;
            else if (sum != 0.0) add_coh(&rv[slot], sum);
          }
        }
        if (SOLO) {
          __syncthreads();
        } else {
          nbar++;
          grid_sync(a.bar, nbar * (unsigned)G, a.err);
        }
        if (tid == 0) {
          scal[3] = sc * sqrt(SOLO ? rv[SLOT_SCALED] : ld_coh(&rv[SLOT_SCALED]));
          scal[5] = sc;
        }
        if (tid < nk) vals[tid] = SOLO ? rv[SLOT_SDOT + tid] : ld_coh(&rv[SLOT_SDOT + tid]);
      } else if (tid == 0) {
        scal[3] = sqrt(scal[3]);
        scal[5] = 1.0;
      }
      __syncthreads();
    }
    if (tid == 0 && j < nb) {
      const double alpha = rowj[0], xnorm = scal[3];
      double beta = alpha, tau = 0.0, scale = 0.0, dscale = 0.0;
      if (xnorm != 0.0) {
        const double aa = fabs(alpha);
        const double w = fmax(aa, xnorm), z = fmin(aa, xnorm);
        const double nrm = (z == 0.0) ? w : w * sqrt(1.0 + (z / w) * (z / w));
        beta = -copysign(nrm, alpha);
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
        dscale = scal[5] / (alpha - beta);  // the dots are of the column divided by scal[5]
      }
      scal[0] = beta;
      scal[1] = tau;
      scal[2] = scale;
      scal[6] = dscale;
      scal[7] = 0.0;
      tauL[j] = tau;
    }
    __syncthreads();
    if (j < nb && tid > j && tid < nb) {
      const int k = tid;
      wk[k] = scal[1] * (rowj[k - j] + vals[k - j - 1] * scal[6]);
    }
    if (j >= 1 && tid < j - 1) Gm[tid + PNB * (j - 1)] = vals[64 + tid];
    __syncthreads();

    // ---- apply the reflector to the slab ----
    if (j < nb && scal[1] != 0.0) {
      const double scale = scal[2];
      for (int lr = tid; lr < nloc; lr += PNT) {
        const int r = r0 + lr;
        if (r > j) {
          const double v = slab[lr + j * RS] * scale;
          slab[lr + j * RS] = v;
          for (int k = j + 1; k < nb; k++) slab[lr + k * RS] -= v * wk[k];
        } else if (r == j) {
          slab[lr + j * RS] = scal[0];
          for (int k = j + 1; k < nb; k++) slab[lr + k * RS] -= wk[k];
        }
      }
    }
    __syncthreads();
  }

  // ---- write V and R back ----
  if (!GLOB) store_slab(slab, RS, nb, nloc, a.A + r0, a.lda, nullptr);

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
    const double *v1 = a.scr + OFF_V1;
    for (int t = tid; t < nb * nb; t += PNT) {
      const int r = t % nb, c = t / nb;
      Xl[r + PNB * c] = SOLO ? (r > c ? slab[r + c * RS] : 0.0) : ld_coh(&v1[t]);
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
        const double v = (r == p) ? 1.0 : slab[lr + p * RS];
        s -= v * Xl[p + PNB * c];
      }
      a.Q[(size_t)r + (size_t)c * a.ldq] = s;
    }
  }
}

// last workgroup out re-zeroes the accumulation buffers and the counters
__device__ void exit_protocol(const PanelArgs &a) {
  __syncthreads();
  __shared__ unsigned last;
  const int tid = threadIdx.x;
  if (tid == 0) {
    const unsigned old = __hip_atomic_fetch_add(&a.cnt[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = (old == (unsigned)a.G - 1) ? 1u : 0u;
  }
  __syncthreads();
  if (last) {
    for (int t = tid; t < 3 * SLOT; t += PNT) st_coh(&a.scr[OFF_RED + t], 0.0);
    for (int t = tid; t < 2 * PNB * PNB; t += PNT) st_coh(&a.scr[OFF_G1 + t], 0.0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) {
      reset_barrier(a.bar);
      __hip_atomic_store(&a.cnt[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

template <bool SOLO>
__global__ __launch_bounds__(PNT) void panel_hh_kernel(PanelArgs a) {
  extern __shared__ double smem[];
  const int g = blockIdx.x;
  const int r0 = g * a.rp;
  const int nloc = max(0, min(a.ml, r0 + a.rp) - r0);
  double *slab = smem, *tail = smem + slab_doubles(a.nb, a.rs);
  load_slab(a, slab, r0, nloc);
  unsigned nbar = 0;
  hh_factor<SOLO>(a, slab, tail, r0, nloc, nbar);
  if (!SOLO) exit_protocol(a);
}

// the same on a panel left in global memory (a.rs == a.lda); LDS holds the
// fixed part only
__global__ __launch_bounds__(PNT) void panel_hh_global_kernel(PanelArgs a) {
  extern __shared__ double smem[];
  const int g = blockIdx.x;
  const int r0 = g * a.rp;
  const int nloc = max(0, min(a.ml, r0 + a.rp) - r0);
  unsigned nbar = 0;
  hh_factor<false, true>(a, a.A + r0, smem, r0, nloc, nbar);
  exit_protocol(a);
}

// ---------------------------------------------------------------------------
// Cholesky-QR2 + Householder reconstruction
// ---------------------------------------------------------------------------

// Gram X^T X of the slab (nloc x nb, stride RS) into M (ld LDM) on MFMA:
// wave w owns the 16x16 output tile (w & 1, w >> 1); 4 independent
// accumulators over the row dimension.  (A operand lane i + 16 kk carries
// X[k0 + kk][ti*16 + i], B operand lane j + 16 kk carries X[k0 + kk][tj*16 + j];
// result lane l, register v holds D[(l/16) + 4v][l%16].)
__device__ void slab_gram(const double *X, int RS, int nloc, int nb, double *M) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ti = wave & 1, tj = wave >> 1;
  const int ci = ti * 16 + (lane & 15), cj = tj * 16 + (lane & 15);
  d4_t acc[4];
#pragma unroll
  for (int u = 0; u < 4; u++) acc[u] = d4_t{0.0, 0.0, 0.0, 0.0};
  const bool vi = ci < nb, vj = cj < nb;
  int k0 = 0;
  for (; k0 + 16 <= nloc; k0 += 16) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int k = k0 + 4 * u + (lane >> 4);
      const double x = vi ? X[k + ci * RS] : 0.0;
      const double y = vj ? X[k + cj * RS] : 0.0;
      acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc[u], 0, 0, 0);
    }
  }
  for (; k0 < nloc; k0 += 4) {
    const int k = k0 + (lane >> 4);
    const double x = (vi && k < nloc) ? X[k + ci * RS] : 0.0;
    const double y = (vj && k < nloc) ? X[k + cj * RS] : 0.0;
    acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc[0], 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const double d = (acc[0][v] + acc[1][v]) + (acc[2][v] + acc[3][v]);
    const int r = ti * 16 + (lane >> 4) + 4 * v, c = tj * 16 + (lane & 15);
    M[r * LDM + c] = d;
  }
}

// X (nloc rows of the slab) <- (X diag(s)) R^-1 for upper R (32 x 32, ld
// LDM, identity beyond nb) with D = 1 / diag(R); s == nullptr means no
// scaling.  R^-1 is formed in MI by one wave, then every wave multiplies
// 16-row tiles of the slab by diag(s) R^-1 on v_mfma_f64_16x16x4f64 (R^-1
// upper: the first output column block takes half the K steps) and writes
// them back in place (a wave reads all of its tile's rows before writing).
__device__ __forceinline__ void slab_mul(double *X, int RS, int nloc, int nb, const double *s,
                                         const double *MI) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ntile = (nloc + 15) / 16, cl = lane & 15;
  for (int t = wave; t < ntile; t += PNT / 64) {
    const int row = 16 * t + cl;
    d4_t acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < PNB / 4; ks++) {
      const int k = 4 * ks + (lane >> 4);
      const double a = (row < nloc && k < nb) ? X[row + k * RS] : 0.0;
      const double sk = (s && k < nb) ? s[k] : 1.0;  // s[k >= nb] is not set
      if (ks < PNB / 8)
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sk * MI[k * LDM + cl], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sk * MI[k * LDM + 16 + cl], acc1, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = 16 * t + (lane >> 4) + 4 * v;
      if (r < nloc) {
        if (cl < nb) X[r + cl * RS] = acc0[v];
        if (16 + cl < nb) X[r + (16 + cl) * RS] = acc1[v];
      }
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void slab_mul_inv(double *X, int RS, int nloc, int nb, const double *R,
                                             const double *D, const double *s, double *MI) {
  if (threadIdx.x < 64) small32::inv_upper32(R, D, MI);
  __syncthreads();
  slab_mul(X, RS, nloc, nb, s, MI);
}

// the reduced 32 x 32 Gram (upper part, first nb) into M (row-major, ld
// LDM), identity beyond nb, zero below the diagonal; all of a thread's
// loads are issued before the first is used (one round trip to the
// coherence point, not one per entry)
__device__ void gram32_read(const double *G, int nb, double *M) {
  constexpr int PER = PNB * PNB / PNT;
  double g[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) g[u] = ld_coh(&G[threadIdx.x + u * PNT]);
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int t = threadIdx.x + u * PNT, r = t / PNB, c = t % PNB;
    M[r * LDM + c] = (r < nb && c < nb) ? (r <= c ? g[u] : 0.0) : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
}

// The Cholesky factor of a Gram G = I + E that lies within NEAR_ID of the
// identity entrywise, without the column-by-column chain: with X the strict
// upper part of E plus half its diagonal, R = I + X gives R^T R = G + X^T X,
// and R^-1 = (I + X)^-1 = I - X + X^2 - ...  For n <= 64 columns ||X||_F <=
// ||E||_F <= 64 NEAR_ID < 1e-9, so the residual X^T X and the dropped X^2
// are below 1e-18, a hundredth of the rounding unit: R is the Cholesky
// factor and I - X its inverse to working precision.  This is the second
// Cholesky-QR pass of a panel whose first pass worked (G = Q1^T Q1 with
// ||I - G|| ~ kappa^2 u); a Gram further from I takes the Cholesky routine.
constexpr double NEAR_ID = 1e-11;
__device__ __forceinline__ bool near_id(double g, bool diag) {
  return fabs(diag ? g - 1.0 : g) <= NEAR_ID;  // false for NaN
}

// 32 x 32 form (row-major, ld LDM, upper part read, identity beyond the
// panel width): R into M (zero below the diagonal), R^-1 into MI.  All
// threads; returns false and writes nothing when G is not near I.
__device__ bool near_identity32(double *M, double *MI, int *flagp) {
  const int tid = threadIdx.x;
  if (tid == 0) *flagp = 1;
  __syncthreads();
  for (int t = tid; t < PNB * PNB; t += PNT) {
    const int r = t / PNB, c = t % PNB;
    if (r <= c && !near_id(M[r * LDM + c], r == c)) *flagp = 0;
  }
  __syncthreads();
  const bool ok = *flagp != 0;
  if (ok) {
    for (int t = tid; t < PNB * PNB; t += PNT) {
      const int r = t / PNB, c = t % PNB;
      const double g = M[r * LDM + c];
      M[r * LDM + c] = (r < c) ? g : (r == c ? 0.5 * (1.0 + g) : 0.0);
      MI[r * LDM + c] = (r < c) ? -g : (r == c ? 0.5 * (3.0 - g) : 0.0);
    }
  }
  __syncthreads();
  return ok;
}

// A square panel (ml == nb): its last column has no rows below the
// diagonal, and dgeqr3's dlarfg returns tau = 0 for it (H = I; R's entry
// keeps the sign it has).  The LU (row-major M) ran over the first nb - 1
// columns: column j = nb - 1 was eliminated by them, not itself; its
// diagonal holds the Schur complement t, +-1 up to rounding since H_j is
// orthogonal, and its U entries above carry no sign yet.  S_j = sign(t)
// makes the pivot S_j t - 1 = -tau_j zero: the pivot is set to exactly 0
// (D_j to 0) and the column's U entries take the sign.  False (nothing
// changed) when |t| is not within 1e-8 of 1.  Every thread returns the same.
__device__ bool square_last32(double *M, double *S, double *D, int nb) {
  const int tid = threadIdx.x, j = nb - 1;
  const double t = M[j * LDM + j];
  const bool good = fabs(fabs(t) - 1.0) <= 1e-8;
  const double s = (t > 0.0) ? 1.0 : -1.0;
  __syncthreads();  // t read everywhere before it changes
  if (good) {
    if (tid < j) M[tid * LDM + j] *= s;
    if (tid == j) {
      M[j * LDM + j] = 0.0;
      S[j] = s;
      D[j] = 0.0;
    }
  }
  __syncthreads();
  return good;
}

// the fallback path out of line, so that its registers do not weigh on the
// Cholesky-QR path's allocation
__device__ __noinline__ void hh_fallback(const PanelArgs &a, double *slab, double *tail, int r0,
                                         int nloc, unsigned &nbar) {
  hh_factor<false>(a, slab, tail, r0, nloc, nbar);
}

__global__ __launch_bounds__(PNT) __attribute__((amdgpu_waves_per_eu(1, 1))) void panel_cqr_kernel(PanelArgs a) {
  extern __shared__ double smem[];
  const int nb = a.nb, RS = a.rs, G = a.G;
  const int g = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int r0 = g * a.rp;
  const int nloc = max(0, min(a.ml, r0 + a.rp) - r0);
  double *X = smem;
  double *tail = smem + slab_doubles(nb, RS);
  double *M0 = tail, *M2 = M0 + PNB * LDM, *M3 = M2 + PNB * LDM;
  double *M4 = M3 + PNB * LDM;  // inverse of the current triangular factor
  double *Sv = M4 + PNB * LDM;  // 32 signs
  double *D0 = Sv + PNB;        // 1 / diag of the current triangular factor
  double *D2 = D0 + PNB;
  int *flag = (int *)(D2 + PNB);  // [0] proceed, [1] the LU's verdict
  double *G1 = a.scr + OFF_G1, *G2 = a.scr + OFF_G2;
  unsigned nbar = 0;
  bool near2 = false;  // pass 2's Gram within NEAR_ID of I: no Cholesky chain

  // the panel's top nb rows (X_top, column-major in M3 with stride LDM) go
  // through both passes beside the slab, with the same arithmetic as those
  // rows get in workgroup 0's slab, so that every workgroup forms the same
  // Q_top -- and the same sign-chosen LU -- bit for bit, with no exchange
  for (int t = tid; t < nb * nb; t += PNT) {
    const int r = t % nb, c = t / nb;
    M3[r + c * LDM] = (r < a.ml) ? a.A[(size_t)r + (size_t)c * a.lda] : 0.0;
  }
  load_slab(a, X, r0, nloc);

  // ---- pass 1: G1 = X^T X, R1 = chol(G1), X <- X R1^-1 ----
  slab_gram(X, RS, nloc, nb, M0);
  __syncthreads();
  for (int t = tid; t < nb * nb; t += PNT) {
    const int r = t / nb, c = t % nb;
    if (r <= c) add_coh(&G1[r * PNB + c], M0[r * LDM + c]);
  }
  nbar++;
  grid_sync(a.bar, nbar * (unsigned)G, a.err);
  gram32_read(G1, nb, M0);
  if (wave == 0) {
    const bool ok = wave_chol32(M0, D0, nb);
    if (tid == 0) flag[0] = ok ? 1 : 0;
  }
  __syncthreads();
  if (flag[0]) {
    slab_mul_inv(X, RS, nloc, nb, M0, D0, nullptr, M4);
    slab_mul(M3, LDM, nb, nb, nullptr, M4);  // X_top R1^-1
    // ---- pass 2 ----
    slab_gram(X, RS, nloc, nb, M2);
    __syncthreads();
    for (int t = tid; t < nb * nb; t += PNT) {
      const int r = t / nb, c = t % nb;
      if (r <= c) add_coh(&G2[r * PNB + c], M2[r * LDM + c]);
    }
    nbar++;
    grid_sync(a.bar, nbar * (unsigned)G, a.err);
    gram32_read(G2, nb, M2);
    near2 = near_identity32(M2, M4, &flag[3]);  // R2 and R2^-1 without the chain
    if (!near2) {
      if (wave == 0) {
        const bool ok = wave_chol32(M2, D2, nb);
        if (tid == 0) flag[0] = ok ? 1 : 0;
      }
      __syncthreads();
    }
  }
  if (flag[0]) {
    if (near2) slab_mul(X, RS, nloc, nb, nullptr, M4);      // X = Q
    else slab_mul_inv(X, RS, nloc, nb, M2, D2, nullptr, M4);
    slab_mul(M3, LDM, nb, nb, nullptr, M4);                 // X_top = Q_top
    if (g == 0) {
      // R = R2 R1 (both upper) on MFMA, one 16 x 16 tile per wave, into M0
      const int ti = wave & 1, tj = wave >> 1;
      d4_t racc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < PNB / 4; ks++) {
        const int k = 4 * ks + (lane >> 4);
        racc = __builtin_amdgcn_mfma_f64_16x16x4f64(M2[(ti * 16 + (lane & 15)) * LDM + k],
                                                    M0[k * LDM + tj * 16 + (lane & 15)], racc, 0, 0, 0);
      }
      __syncthreads();
#pragma unroll
      for (int v = 0; v < 4; v++) M0[(ti * 16 + (lane >> 4) + 4 * v) * LDM + tj * 16 + (lane & 15)] = racc[v];
    }
    __syncthreads();
    // Q_top into M2 (row-major, identity beyond nb) and its sign-chosen LU
    // in every workgroup: U on and above the diagonal, L below, S, 1 / pivots
    for (int t = tid; t < PNB * PNB; t += PNT) {
      const int r = t / PNB, c = t % PNB;
      M2[r * LDM + c] = (r < nb && c < nb) ? M3[r + c * LDM] : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    const bool square = a.ml == nb;
    if (wave == 0) {
      const bool ok = wave_lu_sign32(M2, Sv, D2, square ? nb - 1 : nb);
      if (tid == 0) flag[0] = ok ? 1 : 0;
    }
    __syncthreads();
    if (square && flag[0]) {
      const bool good = square_last32(M2, Sv, D2, nb);
      if (tid == 0 && !good) flag[0] = 0;
      __syncthreads();
    }
  }

  if (!flag[0]) {
    // fall back to the column-by-column Householder on the original panel
    load_slab(a, X, r0, nloc);
    hh_fallback(a, X, tail, r0, nloc, nbar);
    exit_protocol(a);
    return;
  }

  // ---- outputs ----
  // explicit panel Q = Q S
  if (a.Q) store_slab(X, RS, nb, nloc, a.Q + r0, a.ldq, Sv);
  __syncthreads();
  // V2 = (Q2 S) U^-1 (the top rows of workgroup 0 are replaced below)
  slab_mul_inv(X, RS, nloc, nb, M2, D2, Sv, M4);
  __syncthreads();
  if (g == 0) {
    // top nb rows: R' = S R2 R1 on and above the diagonal, V1 = L below
    for (int t = tid; t < nb * nb; t += PNT) {
      const int r = t % nb, c = t / nb;
      X[r + c * RS] = (r <= c) ? Sv[r] * M0[r * LDM + c] : M2[r * LDM + c];
    }
    // T = -U V1^-T = -U (V1^T)^-1: V1^T (unit upper, identity beyond nb)
    // into M3, its inverse into M4 by one wave, then the product of the two
    // upper triangles on MFMA
    for (int t = tid; t < PNB * PNB; t += PNT) {
      const int r = t / PNB, c = t % PNB;
      M3[r * LDM + c] = (r < c) ? M2[c * LDM + r] : (r == c ? 1.0 : 0.0);
    }
    if (tid < PNB) D0[tid] = 1.0;
    __syncthreads();
    if (tid < 64) small32::inv_upper32(M3, D0, M4);
    __syncthreads();
    const int ti = wave & 1, tj = wave >> 1;
    d4_t tacc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < PNB / 4; ks++) {
      const int k = 4 * ks + (lane >> 4), i = ti * 16 + (lane & 15);
      const double u = (i <= k) ? M2[i * LDM + k] : 0.0;  // U; L lies below
      tacc = __builtin_amdgcn_mfma_f64_16x16x4f64(u, M4[k * LDM + tj * 16 + (lane & 15)], tacc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = ti * 16 + (lane >> 4) + 4 * v, c = tj * 16 + (lane & 15);
      // a square panel's last reflector is H = I: its T column is exactly 0
      if (r <= c && c < nb) a.T[(size_t)r + (size_t)c * a.ldt] = (a.ml == nb && c == nb - 1) ? 0.0 : -tacc[v];
    }
  }
  __syncthreads();
  store_slab(X, RS, nb, nloc, a.A + r0, a.lda, nullptr);
  exit_protocol(a);
}

// ---------------------------------------------------------------------------
// 64-column Cholesky-QR2 + Householder reconstruction: the algorithm of
// panel_cqr_kernel on an ml x nb panel with 32 < nb <= 64 (dgeqr3's two
// 32-column panels, the update between them and the T join in one launch).
// The 64 x 64 triangular factors are 2 x 2 grids of 32 x 32 blocks (ld LDM)
// handled by the one-wave 32 x 32 routines of small32.h and 32 x 32 x 32
// MFMA products by all four waves:
//   chol64: R11 = chol(A11), R12 = R11^-T A12, R22 = chol(A22 - R12^T R12);
//   inverse: [X11 X12; 0 X22], X12 = -X11 R12 X22;
//   sign-chosen LU of Q_top (B = Q S - I, the column rule decided on the
//   fly): LU of Q11 (signs S1), L21 = (Q21 S1) U11^-1, Y = L11^-1 Q12, LU of
//   Q22 - L21 Y (signs S2), U12 = Y S2.
// Upper-triangular 64 x 64 matrices are stored as three blocks (11, 12, 22),
// the LU as four.  A panel that fails an acceptance test is factored by the
// column-by-column loop in the same launch: its left 32 columns, their block
// reflector applied to the right ones (one grid reduction of V1^T A2), the
// right columns below row 32, and the T join (a grid reduction of V1^T V2).
// ---------------------------------------------------------------------------
constexpr int B32 = PNB * LDM;                  // doubles of one 32 x 32 block
constexpr int OFF64_G1 = SCRATCH;               // 64 x 64 Gram, pass 1
constexpr int OFF64_G2 = OFF64_G1 + 4096;       // pass 2
constexpr int OFF64_W = OFF64_G2 + 4096;        // fallback: V1^T A2 (32 x 32)
constexpr int OFF64_H = OFF64_W + 1024;         // fallback: V1^T V2 (32 x 32)
constexpr int SCRATCH64 = OFF64_H + 1024;
// the grid barrier's counters: 1 + BAR_GROUPS unsigned ints, 4 KiB apart
constexpr int OFF_BAR = ((SCRATCH64 + 511) / 512) * 512;
constexpr int SCRATCH_ALL = OFF_BAR + 512 * (BAR_GROUPS + 1);
// LDS past the slab: M0 (3 blocks), M2 (4), M4 (3), S, D0, D2, ones (64
// each), flags
constexpr int CQR64_TAIL = 10 * B32 + 4 * 64 + 4;
static_assert(HH_TAIL <= 7 * B32, "the column loop's LDS inside M2 and M4");

// block (bi, bj), bi <= bj, of an upper 64 x 64 matrix stored as 3 blocks
__device__ __forceinline__ double *ublk(double *M, int bi, int bj) { return M + (bi + bj) * B32; }
// block (bi, bj) of a full 64 x 64 matrix stored as 4 blocks
__device__ __forceinline__ double *fblk(double *M, int bi, int bj) { return M + (2 * bi + bj) * B32; }

// entry (r, c) of an upper 64 x 64 matrix in 3-block storage (0 below)
__device__ __forceinline__ double uval(const double *M, int r, int c) {
  if (r >= 32 && c < 32) return 0.0;
  return M[((r >> 5) + (c >> 5)) * B32 + (r & 31) * LDM + (c & 31)];
}
// U(r, c) of the LU in 4-block storage: its upper part (L lies below)
__device__ __forceinline__ double luu(const double *M, int r, int c) {
  if (r > c) return 0.0;
  return M[(2 * (r >> 5) + (c >> 5)) * B32 + (r & 31) * LDM + (c & 31)];
}
// L(r, c), r > c, of the LU in 4-block storage (strict lower part)
__device__ __forceinline__ double lul(const double *M, int r, int c) {
  return M[(2 * (r >> 5) + (c >> 5)) * B32 + (r & 31) * LDM + (c & 31)];
}


// upper Cholesky of the 64 x 64 G in M (3-block storage, upper part read,
// identity beyond nb) in place; 1 / diag(R) into D; Sc a free block.
// Returns false when a pivot is not safely positive or R's diagonal over
// the first nb columns spreads by more than 1e5.
__device__ bool wg_chol64(double *M, double *D, double *Sc, int nb, int *flag) {
  const int tid = threadIdx.x, wave = tid >> 6;
  if (wave == 0) {
    const bool ok = wave_chol32(ublk(M, 0, 0), D, nb < 32 ? nb : 32);
    if (tid == 0) flag[0] = ok ? 1 : 0;
  }
  __syncthreads();
  if (wave == 0) small32::inv_upper32(ublk(M, 0, 0), D, Sc);  // R11^-1
  __syncthreads();
  // R12 = R11^-T A12, then A22 -= R12^T R12
  wg_mm32(Sc, 1, LDM, false, ublk(M, 0, 1), LDM, 1, ublk(M, 0, 1), 1.0, false);
  wg_mm32(ublk(M, 0, 1), 1, LDM, false, ublk(M, 0, 1), LDM, 1, ublk(M, 1, 1), -1.0, true);
  if (wave == 0) {
    const bool ok = wave_chol32(ublk(M, 1, 1), D + 32, nb > 32 ? nb - 32 : 0);
    if (tid == 0) flag[0] = (flag[0] && ok) ? 1 : 0;
  }
  __syncthreads();
  if (wave == 0) {
    // R's diagonal spread over the first nb columns: max/min of diag(R) is
    // max/min of D = 1 / diag(R)
    const int lane = tid & 63;
    double dmn = (lane < nb) ? D[lane] : 1e308, dmx = (lane < nb) ? D[lane] : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      dmn = fmin(dmn, __shfl_xor(dmn, o));
      dmx = fmax(dmx, __shfl_xor(dmx, o));
    }
    if (lane == 0 && !(dmn > 1e-5 * dmx)) flag[0] = 0;
  }
  __syncthreads();
  return flag[0] != 0;
}

// gram64_read and near_identity32's test and factor in one pass for the
// 64 x 64 Gram: from the reduced G (upper, first nb; identity beyond), R =
// I + X into M and R^-1 = I - X into MI (3-block storage, zero below the
// diagonal) when every entry lies within NEAR_ID of the identity's;
// otherwise G into M as gram64_read does.  Returns whether G was near I.
// *flagp (LDS) must be 0 on entry, set before a barrier every thread passed.
__device__ bool gram64_read_near(const double *G, int nb, double *M, double *MI, int *flagp) {
  constexpr int PER = 64 * 64 / PNT;
  double g[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) g[u] = ld_coh(&G[threadIdx.x + u * PNT]);
  int bad = 0;
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int t = threadIdx.x + u * PNT, r = t >> 6, c = t & 63;
    g[u] = (r < nb && c < nb) ? (r <= c ? g[u] : 0.0) : (r == c ? 1.0 : 0.0);
    if (r <= c && !near_id(g[u], r == c)) bad = 1;
  }
  if (bad) *flagp = 1;
  __syncthreads();
  const bool near = *flagp == 0;
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int t = threadIdx.x + u * PNT, r = t >> 6, c = t & 63;
    if (r >= 32 && c < 32) continue;
    const int o = ((r >> 5) + (c >> 5)) * B32 + (r & 31) * LDM + (c & 31);
    if (near) {
      M[o] = (r < c) ? g[u] : (r == c ? 0.5 * (1.0 + g[u]) : 0.0);
      MI[o] = (r < c) ? -g[u] : (r == c ? 0.5 * (3.0 - g[u]) : 0.0);
    } else {
      M[o] = g[u];
    }
  }
  __syncthreads();
  return near;
}

// X = R^-1 for the upper 64 x 64 R given as its blocks R11, R12, R22 (lower
// parts of R11, R22 not read), D = 1 / diag(R); X in 3-block storage
__device__ void wg_inv64(const double *R11, const double *R12, const double *R22, const double *D,
                         double *X) {
  const int wave = threadIdx.x >> 6;
  if (wave == 0) small32::inv_upper32(R11, D, ublk(X, 0, 0));
  else if (wave == 1) small32::inv_upper32(R22, D + 32, ublk(X, 1, 1));
  __syncthreads();
  // X12 = -(X11 R12) X22
  wg_mm32(ublk(X, 0, 0), LDM, 1, false, R12, LDM, 1, ublk(X, 0, 1), 1.0, false);
  wg_mm32(ublk(X, 0, 1), LDM, 1, false, ublk(X, 1, 1), LDM, 1, ublk(X, 0, 1), -1.0, false);
}

// X (nrows rows, stride RS, the first nb of 64 columns) <- (X diag(s)) MI
// for the upper 64 x 64 MI in 3-block storage (s == nullptr: no scaling):
// 16-row tiles on MFMA, a wave reading all of its tile's rows before
// writing them; each output row depends on its own input row only
__device__ void slab_mul64(double *X, int RS, int nrows, int nb, const double *s, double *MI) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, cl = lane & 15;
  const int ntile = (nrows + 15) / 16;
  // this lane's MFMA operands of MI (the same for every tile), diag(s)
  // folded in: (X diag(s)) MI = X (diag(s) MI)
  double mf[16][4];
#pragma unroll
  for (int ks = 0; ks < 16; ks++) {
    const int k = 4 * ks + (lane >> 4);
    const double sk = (s && k < nb) ? s[k] : 1.0;
#pragma unroll
    for (int q = 0; q < 4; q++) mf[ks][q] = (4 * ks <= 16 * q + 15) ? sk * uval(MI, k, 16 * q + cl) : 0.0;
  }
  for (int t = wave; t < ntile; t += PNT / 64) {
    const int row = 16 * t + cl;
    d4_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] = d4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < 16; ks++) {
      const int k = 4 * ks + (lane >> 4);
      const double a = (row < nrows && k < nb) ? X[row + k * RS] : 0.0;
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (4 * ks <= 16 * q + 15)  // MI upper: rows k <= 16 q + 15 only
          acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, mf[ks][q], acc[q], 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = 16 * t + (lane >> 4) + 4 * v;
      if (r < nrows)
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (16 * q + cl < nb) X[r + (16 * q + cl) * RS] = acc[q][v];
    }
  }
  __syncthreads();
}

// Gram X^T X of the slab's first nloc rows (64 columns, upper 16 x 16 tiles
// dealt to the waves), its upper part added into the global G (64 x 64)
__device__ void slab_gram64_add(const double *X, int RS, int nloc, int nb, double *G) {
  if (nloc <= 0) return;  // a workgroup without rows (the one that forms T)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int u = wave; u < 10; u += 4) {
    // tile u of (0,0) (0,1) (0,2) (0,3) (1,1) (1,2) (1,3) (2,2) (2,3) (3,3)
    const int ti = (u < 4) ? 0 : (u < 7) ? 1 : (u < 9) ? 2 : 3;
    const int tj = (u < 4) ? u : (u < 7) ? u - 3 : (u < 9) ? u - 5 : 3;
    const int ci = ti * 16 + (lane & 15), cj = tj * 16 + (lane & 15);
    const bool vi = ci < nb, vj = cj < nb;
    d4_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] = d4_t{0.0, 0.0, 0.0, 0.0};
    int k0 = 0;
    for (; k0 + 16 <= nloc; k0 += 16) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int k = k0 + 4 * q + (lane >> 4);
        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(vi ? X[k + ci * RS] : 0.0, vj ? X[k + cj * RS] : 0.0,
                                                      acc[q], 0, 0, 0);
      }
    }
    for (; k0 < nloc; k0 += 4) {
      const int k = k0 + (lane >> 4);
      acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64((vi && k < nloc) ? X[k + ci * RS] : 0.0,
                                                    (vj && k < nloc) ? X[k + cj * RS] : 0.0, acc[0], 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = ti * 16 + (lane >> 4) + 4 * v, c = tj * 16 + (lane & 15);
      const double d = (acc[0][v] + acc[1][v]) + (acc[2][v] + acc[3][v]);
      if (r <= c && c < nb) add_coh(&G[r * 64 + c], d);
    }
  }
}

// the reduced Gram (upper, first nb) into M (3-block storage), identity
// beyond nb, zero below the diagonal
__device__ void gram64_read(const double *G, int nb, double *M) {
  // all of a thread's loads are issued before the first is used (one
  // round trip to the coherence point, not one per entry): every entry of
  // G is loaded, the ones not needed are then dropped
  constexpr int PER = 64 * 64 / PNT;
  double g[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) g[u] = ld_coh(&G[threadIdx.x + u * PNT]);
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int t = threadIdx.x + u * PNT, r = t >> 6, c = t & 63;
    if (r >= 32 && c < 32) continue;
    const double v = (r < nb && c < nb) ? (r <= c ? g[u] : 0.0) : (r == c ? 1.0 : 0.0);
    M[((r >> 5) + (c >> 5)) * B32 + (r & 31) * LDM + (c & 31)] = v;
  }
  __syncthreads();
}

// square_last32 for the LU in 4-block storage (the 64-column panel)
__device__ bool square_last64(double *M, double *S, double *D, int nb) {
  const int tid = threadIdx.x, j = nb - 1;
  auto at = [M](int r, int c) -> double & {
    return M[(2 * (r >> 5) + (c >> 5)) * B32 + (r & 31) * LDM + (c & 31)];
  };
  const double t = at(j, j);
  const bool good = fabs(fabs(t) - 1.0) <= 1e-8;
  const double s = (t > 0.0) ? 1.0 : -1.0;
  __syncthreads();  // t read everywhere before it changes
  if (good) {
    if (tid < j) at(tid, j) *= s;
    if (tid == j) {
      at(j, j) = 0.0;
      S[j] = s;
      D[j] = 0.0;
    }
  }
  __syncthreads();
  return good;
}

// sign-chosen LU of Q_top (4-block storage in M, identity beyond nb) in
// place: U on and above the diagonal, L below; S the signs, D 1 / pivots.
// Sc: 3 free blocks, one1: 32 ones.  Returns false when some |t| > 1 -
// 1e-10 (tau within 1e-10 of 0 or 2).
__device__ bool wg_lu64(double *M, double *S, double *D, double *Sc, const double *one1, int nb, int *flag) {
  const int tid = threadIdx.x, wave = tid >> 6;
  double *Q11 = fblk(M, 0, 0), *Q12 = fblk(M, 0, 1), *Q21 = fblk(M, 1, 0), *Q22 = fblk(M, 1, 1);
  double *Ui = Sc, *Lt = Sc + B32, *Lti = Sc + 2 * B32;
  if (wave == 0) {
    const bool ok = wave_lu_sign32(Q11, S, D, nb < 32 ? nb : 32);
    if (tid == 0) flag[0] = ok ? 1 : 0;
  }
  __syncthreads();
  // U11^-1 (wave 0) and (L11^T)^-1 = (L11^-1)^T (wave 1); Q21 S1
  for (int t = tid; t < PNB * PNB; t += PNT) {
    const int r = t / PNB, c = t % PNB;
    Lt[r * LDM + c] = (r < c) ? Q11[c * LDM + r] : (r == c ? 1.0 : 0.0);
    Q21[r * LDM + c] *= S[c];
  }
  __syncthreads();
  if (wave == 0) small32::inv_upper32(Q11, D, Ui);
  else if (wave == 1) small32::inv_upper32(Lt, one1, Lti);
  __syncthreads();
  wg_mm32(Q21, LDM, 1, false, Ui, LDM, 1, Q21, 1.0, false);    // L21 = (Q21 S1) U11^-1
  wg_mm32(Lti, 1, LDM, false, Q12, LDM, 1, Q12, 1.0, false);   // Y = L11^-1 Q12
  wg_mm32(Q21, LDM, 1, false, Q12, LDM, 1, Q22, -1.0, true);   // Q22 - L21 Y
  if (wave == 0) {
    const bool ok = wave_lu_sign32(Q22, S + 32, D + 32, nb > 32 ? nb - 32 : 0);
    if (tid == 0) flag[0] = (flag[0] && ok) ? 1 : 0;
  }
  __syncthreads();
  for (int t = tid; t < PNB * PNB; t += PNT) {
    const int r = t / PNB, c = t % PNB;
    Q12[r * LDM + c] *= S[32 + c];  // U12 = Y S2
  }
  __syncthreads();
  return flag[0] != 0;
}

// last workgroup out re-zeroes the 64-wide accumulation buffers and the
// counters
// The 64-wide panel keeps no exit protocol: its launches alternate between
// two scratch sets (accumulators, reduction slots, barrier counters), and
// each launch clears the set the next one will use -- the previous launch
// on the stream, which used it, has finished -- from a workgroup that has
// time for it, so no workgroup waits on an exit counter at the end.
__device__ void clear_set64(double *scr) {
  for (int t = threadIdx.x; t < 3 * SLOT; t += PNT) st_coh(&scr[OFF_RED + t], 0.0);
  for (int t = threadIdx.x; t < SCRATCH64 - OFF64_G1; t += PNT) st_coh(&scr[OFF64_G1 + t], 0.0);
  if (threadIdx.x == 0) reset_barrier((unsigned *)(scr + OFF_BAR));
}


// V1(r0 + lr, k) of the slab after the left columns' column loop: unit
// lower in the panel's rows 0..31 (R11 lies above), as stored below
__device__ __forceinline__ double v1val(const double *slab, int RS, int r0, int lr, int k) {
  const int r = r0 + lr;
  return (r < k) ? 0.0 : (r == k ? 1.0 : slab[lr + k * RS]);
}

// the column-by-column fallback of the 64-wide panel (slab reloaded, A, T
// untouched so far): left 32 columns, their reflectors on the right ones,
// the right columns below row 32, the T join.  tail: the column loop's LDS
// (HH_TAIL doubles); blk: three free 32 x 32 blocks (T1, W, H).
__device__ __noinline__ void hh_fallback64(const PanelArgs &a, int tail_off, int blk_off, int r0, int nloc,
                                           unsigned &nbar) {
  // (the LDS pointers are formed here from the dynamic LDS array: passed in
  // as generic pointers and offset, the compiler's null checks of their
  // casts produce an instruction it cannot encode)
  extern __shared__ double smem[];
  double *slab = smem, *tail = smem + tail_off, *blk = smem + blk_off;
  double *T1 = blk, *Wl = blk + B32, *Hl = blk + 2 * B32;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = blockIdx.x;
  const int RS = a.rs, G = a.G, nb2 = a.nb - 32;
  PanelArgs a1 = a;
  a1.nb = 32;
  a1.Q = nullptr;
  hh_factor<false>(a1, slab, tail, r0, nloc, nbar);
  // every workgroup holds T1 (the column loop forms it from reduced values)
  const double *Tl = tail + PNB * PNB;
  for (int t = tid; t < PNB * PNB; t += PNT) T1[(t / PNB) * LDM + t % PNB] = Tl[(t / PNB) + PNB * (t % PNB)];
  __syncthreads();
  // ---- W = V1^T A2 over all rows (grid reduction), W' = T1^T W ----
  double *W = a.scr + OFF64_W;
  {
    const int ti = wave & 1, tj = wave >> 1;
    const int ci = ti * 16 + (lane & 15), cj = tj * 16 + (lane & 15);
    d4_t acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < nloc; k0 += 4) {
      const int lr = k0 + (lane >> 4);
      const double x = (lr < nloc) ? v1val(slab, RS, r0, lr, ci) : 0.0;
      const double y = (lr < nloc && cj < nb2) ? slab[lr + (32 + cj) * RS] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = ti * 16 + (lane >> 4) + 4 * v, c = tj * 16 + (lane & 15);
      if (acc[v] != 0.0) add_coh(&W[r * PNB + c], acc[v]);
    }
  }
  nbar++;
  grid_sync(a.bar, nbar * (unsigned)G, a.err);
  for (int t = tid; t < PNB * PNB; t += PNT) Wl[(t / PNB) * LDM + t % PNB] = ld_coh(&W[t]);
  __syncthreads();
  // W' = T1^T W (T1 upper), into Wl
  wg_mm32(T1, 1, LDM, false, Wl, LDM, 1, Wl, 1.0, false);
  // A2 -= V1 W' on this slab's rows
  for (int t = wave; t < (nloc + 15) / 16; t += PNT / 64) {
    const int lr = 16 * t + (lane & 15);
    d4_t acc[2] = {d4_t{0.0, 0.0, 0.0, 0.0}, d4_t{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int ks = 0; ks < 8; ks++) {
      const int k = 4 * ks + (lane >> 4);
      const double x = (lr < nloc) ? v1val(slab, RS, r0, lr, k) : 0.0;
      acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, Wl[k * LDM + (lane & 15)], acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, Wl[k * LDM + 16 + (lane & 15)], acc[1], 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = 16 * t + (lane >> 4) + 4 * v;
      if (r < nloc) {
        if ((lane & 15) < nb2) slab[r + (32 + (lane & 15)) * RS] -= acc[0][v];
        if (16 + (lane & 15) < nb2) slab[r + (48 + (lane & 15)) * RS] -= acc[1][v];
      }
    }
  }
  __syncthreads();
  // R12: rows 0..31 of the right columns, in workgroup 0's slab
  if (g == 0)
    for (int t = tid; t < 32 * nb2; t += PNT) {
      const int r = t % 32, c = t / 32;
      a.A[(size_t)r + (size_t)(32 + c) * a.lda] = slab[r + (32 + c) * RS];
    }
  // ---- the right columns below row 32 by the column loop ----
  // The right columns move to slab columns 0..31 (V1 to 32..63), and
  // workgroup 0's rows up by 32 (its first 32 rows are done), so that the
  // column loop runs on the slab as it stands.  (Handing it an offset slab
  // pointer instead made the compiler emit an address-space check that it
  // cannot encode.)
  const int skip = (g == 0) ? 32 : 0;
  const int nloc2 = nloc - skip, r02 = r0 + skip - 32;  // rows of the right sub-panel
  // (a slab has fewer rows than the workgroup has threads: one row each)
  for (int c0 = 0; c0 < 32; c0 += 4) {
    double lo[4], hi[4];
    const int lr = tid;
    if (lr < nloc2) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        lo[q] = slab[lr + skip + (c0 + q) * RS];
        hi[q] = slab[lr + skip + (32 + c0 + q) * RS];
      }
    }
    __syncthreads();
    if (lr < nloc2) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        slab[lr + (c0 + q) * RS] = hi[q];
        slab[lr + (32 + c0 + q) * RS] = lo[q];
      }
    }
    __syncthreads();
  }
  PanelArgs a2 = a;
  a2.ml = a.ml - 32;
  a2.nb = nb2;
  a2.A = a.A + 32 + (size_t)32 * a.lda;
  a2.T = a.T + 32 + (size_t)32 * a.ldt;
  a2.Q = nullptr;
  hh_factor<false>(a2, slab, tail, r02, nloc2, nbar);
  // ---- T12 = -T1 (V1^T V2) T2 (V1 below row 32, V2 unit lower there) ----
  double *H = a.scr + OFF64_H;
  {
    const int ti = wave & 1, tj = wave >> 1;
    const int ci = ti * 16 + (lane & 15), cj = tj * 16 + (lane & 15);
    d4_t acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < nloc2; k0 += 4) {
      const int lr = k0 + (lane >> 4), rr = r02 + lr;  // row of the right sub-panel
      double y = 0.0;
      if (lr < nloc2 && cj < nb2) y = (rr < cj) ? 0.0 : (rr == cj ? 1.0 : slab[lr + cj * RS]);
      const double x = (lr < nloc2) ? slab[lr + (32 + ci) * RS] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = ti * 16 + (lane >> 4) + 4 * v, c = tj * 16 + (lane & 15);
      if (acc[v] != 0.0) add_coh(&H[r * PNB + c], acc[v]);
    }
  }
  nbar++;
  grid_sync(a.bar, nbar * (unsigned)G, a.err);
  if (g == 0) {
    // T2 (the column loop's T of the right columns, in LDS after it) and H
    const double *T2 = tail + PNB * PNB;  // the column loop's T, T2[r + PNB c]
    for (int t = tid; t < PNB * PNB; t += PNT) {
      const int r = t / PNB, c = t % PNB;
      Hl[r * LDM + c] = ld_coh(&H[t]);
      Wl[r * LDM + c] = (c < nb2) ? T2[r + PNB * c] : 0.0;  // T2 row-major, zero beyond nb2
    }
    __syncthreads();
    wg_mm32(T1, LDM, 1, false, Hl, LDM, 1, Hl, 1.0, false);    // T1 H
    wg_mm32(Hl, LDM, 1, false, Wl, LDM, 1, Hl, -1.0, false);   // -(T1 H) T2
    for (int t = tid; t < 32 * nb2; t += PNT) {
      const int r = t % 32, c = t / 32;
      a.T[(size_t)r + (size_t)(32 + c) * a.ldt] = Hl[r * LDM + c];
    }
  }
}

__global__ __launch_bounds__(PNT) __attribute__((amdgpu_waves_per_eu(1, 1))) void panel_cqr64_kernel(PanelArgs a) {
  extern __shared__ double smem[];
  const int nb = a.nb, RS = a.rs, G = a.G;
  const int g = blockIdx.x, tid = threadIdx.x;
  const int r0 = g * a.rp;
  const int nloc = max(0, min(a.ml, r0 + a.rp) - r0);
  const int ntop = min(64, a.ml);  // rows of X_top
  double *X = smem;                 // rp slab rows, then X_top at rows rp .. rp + 63
  double *tail = smem + slab_doubles(64, RS);
  double *M0 = tail;                // R1, then R (3 blocks)
  double *M2 = M0 + 3 * B32;        // R2 (3-block storage), then the LU (4 blocks)
  double *M4 = M2 + 4 * B32;        // inverses and scratch (3 blocks)
  double *Sv = M4 + 3 * B32;        // 64 signs
  double *D0 = Sv + 64;             // 1 / diag of the current factor
  double *D2 = D0 + 64;
  double *one = D2 + 64;
  int *flag = (int *)(one + 64);
  double *G1 = a.scr + OFF64_G1, *G2 = a.scr + OFF64_G2;
  unsigned nbar = 0;
  bool near2 = false;  // pass 2's Gram within NEAR_ID of I: no Cholesky chain

  if (g == G - 1) clear_set64(a.scr_next);  // (the last workgroup has the fewest rows)
  // slab and, at rows rp .., the panel's top 64 rows (see panel_cqr_kernel)
  {
    // wave w loads columns w, w + 4, ..., lanes down the rows; all sixteen
    // loads in flight before the first LDS write
    const int lane = tid & 63, wave = tid >> 6;
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int c = wave + 4 * u;
      v[u] = (c < nb && lane < ntop) ? a.A[(size_t)lane + (size_t)c * a.lda] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int c = wave + 4 * u;
      if (c < nb) X[a.rp + lane + c * RS] = v[u];
    }
  }
  if (tid < 64) one[tid] = 1.0;
  load_slab(a, X, r0, nloc);

  // ---- pass 1 ----
  slab_gram64_add(X, RS, nloc, nb, G1);
  nbar++;
  grid_sync(a.bar, nbar * (unsigned)G, a.err);
  gram64_read(G1, nb, M0);
  bool ok = wg_chol64(M0, D0, M4, nb, flag);
  if (ok) {
    wg_inv64(ublk(M0, 0, 0), ublk(M0, 0, 1), ublk(M0, 1, 1), D0, M4);
    slab_mul64(X, RS, a.rp + 64, nb, nullptr, M4);
    // ---- pass 2 ----
    slab_gram64_add(X, RS, nloc, nb, G2);
    if (tid == 0) flag[3] = 0;  // gram64_read_near's verdict (the barrier orders it)
    nbar++;
    grid_sync(a.bar, nbar * (unsigned)G, a.err);
    near2 = gram64_read_near(G2, nb, M2, M4, &flag[3]);
    ok = near2 || wg_chol64(M2, D2, M4, nb, flag);
  }
  if (ok) {
    if (!near2) wg_inv64(ublk(M2, 0, 0), ublk(M2, 0, 1), ublk(M2, 1, 1), D2, M4);
    // only the top rows become Q (Q_top = X_top R2^-1, for the LU): the
    // slab keeps X = A R1^-1, and V2 is formed from it below with the
    // combined transform R2^-1 S U^-1 (one pass over the slab, not two)
    slab_mul64(X + a.rp, RS, 64, nb, nullptr, M4);
    // R = R2 R1 (workgroup 0 only), from registers after all reads
    const int lane = tid & 63, wave = tid >> 6;
    const int ti = wave & 1, tj = wave >> 1;
    d4_t racc[3];
    if (g == 0) {
#pragma unroll
      for (int bq = 0; bq < 3; bq++) {
        const int bi = (bq == 2) ? 1 : 0, bj = (bq == 0) ? 0 : 1;
        racc[bq] = d4_t{0.0, 0.0, 0.0, 0.0};
        const int i = 32 * bi + ti * 16 + (lane & 15), j = 32 * bj + tj * 16 + (lane & 15);
        for (int ks = 8 * bi; ks < 8 * bj + 8; ks++) {
          const int k = 4 * ks + (lane >> 4);
          racc[bq] = __builtin_amdgcn_mfma_f64_16x16x4f64(uval(M2, i, k), uval(M0, k, j), racc[bq], 0, 0, 0);
        }
      }
    }
    __syncthreads();
    // Q_top into M2 (4 blocks, identity beyond nb)
    for (int t = tid; t < 64 * 64; t += PNT) {
      const int r = t >> 6, c = t & 63;
      fblk(M2, r >> 5, c >> 5)[(r & 31) * LDM + (c & 31)] =
          (r < nb && c < nb) ? X[a.rp + r + c * RS] : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    if (g == 0) {
      // R into the X_top rows (free now), column-major like the slab
#pragma unroll
      for (int bq = 0; bq < 3; bq++) {
        const int bi = (bq == 2) ? 1 : 0, bj = (bq == 0) ? 0 : 1;
#pragma unroll
        for (int v = 0; v < 4; v++)
          X[a.rp + 32 * bi + ti * 16 + (lane >> 4) + 4 * v + (32 * bj + tj * 16 + (lane & 15)) * RS] = racc[bq][v];
      }
    }
    // the sign-chosen LU of Q_top (M0 free in every workgroup: scratch)
    const bool square = a.ml == nb;
    ok = wg_lu64(M2, Sv, D2, M0, one, square ? nb - 1 : nb, flag);
    if (ok && square) ok = square_last64(M2, Sv, D2, nb);
  }

  if (!ok) {
    load_slab(a, X, r0, nloc);
    hh_fallback64(a, (int)(M2 - smem), (int)(M0 - smem), r0, nloc, nbar);
    return;
  }

  // ---- outputs: V2 = X (R2^-1 S U^-1) = (Q2 S) U^-1 ----
  // (a workgroup without rows -- the extra one that forms T -- skips them)
  if (nloc > 0) {
  wg_inv64(fblk(M2, 0, 0), fblk(M2, 0, 1), fblk(M2, 1, 1), D2, M0);  // U^-1
  {
    // M4 <- M4 diag(S) M0 (upper x diag x upper), from registers after all reads
    const int lane = tid & 63, wave = tid >> 6;
    const int ti = wave & 1, tj = wave >> 1;
    d4_t macc[3];
#pragma unroll
    for (int bq = 0; bq < 3; bq++) {
      const int bi = (bq == 2) ? 1 : 0, bj = (bq == 0) ? 0 : 1;
      macc[bq] = d4_t{0.0, 0.0, 0.0, 0.0};
      const int i = 32 * bi + ti * 16 + (lane & 15), j = 32 * bj + tj * 16 + (lane & 15);
      for (int ks = 8 * bi; ks < 8 * bj + 8; ks++) {
        const int k = 4 * ks + (lane >> 4);
        const double sk = (k < nb) ? Sv[k] : 1.0;
        macc[bq] = __builtin_amdgcn_mfma_f64_16x16x4f64(uval(M4, i, k) * sk, uval(M0, k, j), macc[bq], 0, 0, 0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int bq = 0; bq < 3; bq++) {
      const int bi = (bq == 2) ? 1 : 0, bj = (bq == 0) ? 0 : 1;
#pragma unroll
      for (int v = 0; v < 4; v++)
        ublk(M4, bi, bj)[(ti * 16 + (lane >> 4) + 4 * v) * LDM + tj * 16 + (lane & 15)] = macc[bq][v];
    }
    __syncthreads();
  }
  slab_mul64(X, RS, nloc, nb, nullptr, M4);
  }
  if (g == 0) {
    // top nb rows: R' = S R2 R1 on and above the diagonal (R from the
    // X_top rows), V1 = L below
    for (int t = tid; t < nb * nb; t += PNT) {
      const int r = t % nb, c = t / nb;
      X[r + c * RS] = (r <= c) ? Sv[r] * X[a.rp + r + c * RS] : lul(M2, r, c);
    }
  }
  if (g == G - 1) {
    // (every workgroup holds U and L: the last one -- launched without
    // rows when the chip has a CU to spare, else the one with the fewest
    // -- forms T while the others form V2 and workgroup 0 writes the top
    // rows; with one workgroup, after them)
    __syncthreads();
    // T = -U (V1^T)^-1: V1^T (unit upper, identity beyond nb) into M0's
    // blocks, its inverse into M4, then the product of the two triangles
    for (int t = tid; t < 64 * 64; t += PNT) {
      const int r = t >> 6, c = t & 63;
      if (r >= 32 && c < 32) continue;
      ublk(M0, r >> 5, c >> 5)[(r & 31) * LDM + (c & 31)] = (r < c) ? lul(M2, c, r) : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    wg_inv64(ublk(M0, 0, 0), ublk(M0, 0, 1), ublk(M0, 1, 1), one, M4);
    const int lane = tid & 63, wave = tid >> 6;
    for (int u = wave; u < 10; u += 4) {
      const int ti = (u < 4) ? 0 : (u < 7) ? 1 : (u < 9) ? 2 : 3;
      const int tj = (u < 4) ? u : (u < 7) ? u - 3 : (u < 9) ? u - 5 : 3;
      const int i = ti * 16 + (lane & 15), j = tj * 16 + (lane & 15);
      d4_t acc = {0.0, 0.0, 0.0, 0.0};
      for (int ks = 4 * ti; ks < 4 * tj + 4; ks++) {
        const int k = 4 * ks + (lane >> 4);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(luu(M2, i, k), uval(M4, k, j), acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = ti * 16 + (lane >> 4) + 4 * v, c = tj * 16 + (lane & 15);
        // a square panel's last reflector is H = I: its T column is exactly 0
        if (r <= c && c < nb) a.T[(size_t)r + (size_t)c * a.ldt] = (a.ml == nb && c == nb - 1) ? 0.0 : -acc[v];
      }
    }
  }
  __syncthreads();
  store_slab(X, RS, nb, nloc, a.A + r0, a.lda, nullptr);
}

// ---------------------------------------------------------------------------
// Cholesky-QR2 + Householder reconstruction for a panel taller than the
// workgroups' LDS can hold: the same algorithm as panel_cqr_kernel, with each
// workgroup's rows streamed through LDS in sub-slabs of a.sr rows once per
// pass instead of staying resident -- Gram of A; Gram of A R1^-1; then
// Q S = A (R^-1 S) and V = (Q S) U^-1 written out (R^-1 = R1^-1 R2^-1 and
// U^-1 formed explicitly).  The panel is only read until the last pass, so
// a failed acceptance test falls back to the column-by-column loop on the
// untouched panel in global memory.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(PNT) __attribute__((amdgpu_waves_per_eu(1, 1))) void panel_cqr_stream_kernel(PanelArgs a) {
  extern __shared__ double smem[];
  const int nb = a.nb, RS = a.rs, G = a.G, SR = a.sr;
  const int g = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int r0 = g * a.rp;
  const int nloc = max(0, min(a.ml, r0 + a.rp) - r0);
  double *X = smem;
  double *tail = smem + slab_doubles(nb, RS);
  double *M0 = tail, *M2 = M0 + PNB * LDM, *M3 = M2 + PNB * LDM;
  double *M4 = M3 + PNB * LDM, *M5 = M4 + PNB * LDM;
  double *Sv = M5 + PNB * LDM, *D0 = Sv + PNB, *D2 = D0 + PNB;
  int *flag = (int *)(D2 + PNB);  // [0] proceed, [1] the LU's verdict
  double *G1 = a.scr + OFF_G1, *G2 = a.scr + OFF_G2, *pub = a.scr + OFF_PUB;
  unsigned nbar = 0;

  // Gram of the sub-slabs (transformed by MI when given) into Gd (global)
  auto gram_pass = [&](double *Gd, const double *MI, double *M) {
    for (int s0 = 0; s0 < nloc; s0 += SR) {
      const int n = min(SR, nloc - s0);
      load_slab(a, X, r0 + s0, n);
      if (MI) slab_mul(X, RS, n, nb, nullptr, MI);
      slab_gram(X, RS, n, nb, M);
      __syncthreads();
      for (int t = tid; t < nb * nb; t += PNT) {
        const int r = t / nb, c = t % nb;
        if (r <= c) add_coh(&Gd[r * PNB + c], M[r * LDM + c]);
      }
      __syncthreads();
    }
  };
  // Gram from global into M (identity beyond nb) and its Cholesky
  auto chol_of = [&](const double *Gd, double *M, double *D) {
    gram32_read(Gd, nb, M);
    if (wave == 0) {
      const bool ok = wave_chol32(M, D, nb);
      if (tid == 0) flag[0] = ok ? 1 : 0;
    }
    __syncthreads();
  };
  // upper X Y on MFMA into Z (one 16 x 16 tile per wave; Z may be X or Y)
  auto upper_mul = [&](const double *Xm, const double *Ym, double *Z) {
    const int ti = wave & 1, tj = wave >> 1;
    d4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < PNB / 4; ks++) {
      const int k = 4 * ks + (lane >> 4);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Xm[(ti * 16 + (lane & 15)) * LDM + k],
                                                 Ym[k * LDM + tj * 16 + (lane & 15)], acc, 0, 0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 4; v++) Z[(ti * 16 + (lane >> 4) + 4 * v) * LDM + tj * 16 + (lane & 15)] = acc[v];
    __syncthreads();
  };

  // ---- pass 1: G1 = A^T A, R1 ----
  gram_pass(G1, nullptr, M0);
  nbar++;
  grid_sync(a.bar, nbar * (unsigned)G, a.err);
  chol_of(G1, M0, D0);
  if (flag[0]) {
    // ---- pass 2: G2 = (A R1^-1)^T (A R1^-1), R2 ----
    if (wave == 0) small32::inv_upper32(M0, D0, M4);
    __syncthreads();
    gram_pass(G2, M4, M2);
    nbar++;
    grid_sync(a.bar, nbar * (unsigned)G, a.err);
    chol_of(G2, M2, D2);
  }
  if (flag[0]) {
    // R^-1 = R1^-1 R2^-1 into M4
    if (wave == 0) small32::inv_upper32(M2, D2, M5);
    __syncthreads();
    upper_mul(M4, M5, M4);
    if (g == 0) {
      // top nb x nb of Q = A R^-1 and its sign-chosen LU
      load_slab(a, X, r0, min(nloc, SR));
      slab_mul(X, RS, nb, nb, nullptr, M4);
      for (int t = tid; t < PNB * PNB; t += PNT) {
        const int r = t / PNB, c = t % PNB;
        M3[r * LDM + c] = (r < nb && c < nb) ? X[r + c * RS] : (r == c ? 1.0 : 0.0);
      }
      __syncthreads();
      if (wave == 0) {
        const bool ok = wave_lu_sign32(M3, Sv, D2, nb);
        if (tid == 0) flag[1] = ok ? 1 : 0;
      }
      __syncthreads();
      for (int t = tid; t < PNB * PNB; t += PNT) {
        const int r = t / PNB, c = t % PNB;
        st_coh(&pub[r * PNB + c], r <= c ? M3[r * LDM + c] : 0.0);
      }
      if (tid < nb) st_coh(&pub[PNB * PNB + tid], Sv[tid]);
      if (tid == 0) st_coh(&pub[PNB * PNB + PNB], flag[1] ? 1.0 : 0.0);
    }
    nbar++;
    grid_arrive(a.bar);
    if (g == 0) upper_mul(M2, M0, M0);  // R = R2 R1 while the others wait
    grid_wait(a.bar, nbar * (unsigned)G, a.err);
    if (tid == 0) flag[0] = ld_coh(&pub[PNB * PNB + PNB]) != 0.0 ? 1 : 0;
    __syncthreads();
  }

  if (!flag[0] && a.decline) {
    // the panel is untouched: report and leave it to the caller's TSQR
    if (blockIdx.x == 0 && tid == 0)
      __hip_atomic_store(&a.err[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    exit_protocol(a);
    return;
  }
  if (!flag[0]) {
    // the panel is untouched: column-by-column Householder on it in global
    // memory (every workgroup took the same decisions from the same data)
    PanelArgs b = a;
    b.rs = a.lda;
    hh_factor<false, true>(b, a.A + r0, smem, r0, nloc, nbar);
    exit_protocol(a);
    return;
  }

  // ---- U, S; R^-1 S into M4, U^-1 into M5 ----
  if (tid < nb) Sv[tid] = ld_coh(&pub[PNB * PNB + tid]);
  for (int t = tid; t < PNB * PNB; t += PNT) {
    const int r = t / PNB, c = t % PNB;
    M2[r * LDM + c] = ld_coh(&pub[r * PNB + c]);
  }
  __syncthreads();
  if (tid < PNB) D2[tid] = 1.0 / M2[tid * LDM + tid];
  if (tid < PNB && tid >= nb) Sv[tid] = 1.0;  // S beyond nb: 1
  __syncthreads();
  if (wave == 0) small32::inv_upper32(M2, D2, M5);
  for (int t = tid; t < PNB * PNB; t += PNT) {
    const int r = t / PNB, c = t % PNB;
    M4[r * LDM + c] *= Sv[c];
  }
  __syncthreads();
  if (g == 0 && tid < nb) {
    // T = -U V1^-T, one row per thread (as in panel_cqr_kernel)
    const int r = tid;
    double t[PNB];
#pragma unroll
    for (int c = 0; c < PNB; c++) {
      const double u = -M2[r * LDM + c];
      t[c] = (c >= r) ? u : 0.0;
    }
#pragma unroll
    for (int q = 0; q < PNB; q++) {
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int c = q + 1; c < PNB; c++) t[c] -= t[q] * M3[c * LDM + q];
    }
#pragma unroll
    for (int c = PNB - 1; c >= 0; c--)
      if (c >= r) a.T[(size_t)r + (size_t)min(c, nb - 1) * a.ldt] = t[c];
  }
  // ---- pass 3: Q S = A (R^-1 S), V = (Q S) U^-1 ----
  for (int s0 = 0; s0 < nloc; s0 += SR) {
    const int n = min(SR, nloc - s0);
    load_slab(a, X, r0 + s0, n);
    slab_mul(X, RS, n, nb, nullptr, M4);
    if (a.Q) store_slab(X, RS, nb, n, a.Q + r0 + s0, a.ldq, nullptr);
    __syncthreads();
    slab_mul(X, RS, n, nb, nullptr, M5);
    if (g == 0 && s0 == 0) {
      // top nb rows: R' = S R2 R1 on and above the diagonal, V1 = L below
      for (int t = tid; t < nb * nb; t += PNT) {
        const int r = t % nb, c = t / nb;
        X[r + c * RS] = (r <= c) ? Sv[r] * M0[r * LDM + c] : M3[r * LDM + c];
      }
      __syncthreads();
    }
    store_slab(X, RS, nb, n, a.A + r0 + s0, a.lda, nullptr);
    __syncthreads();
  }
  exit_protocol(a);
}

// ---------------------------------------------------------------------------
// TSQR (tall panels): a tree of independent slab QRs.  Level L splits its
// M x nb matrix into S slabs of nearly equal height (slab s: rows
// [s M / S, (s + 1) M / S), each >= nb); workgroup s factors its slab in LDS
// with the column-by-column Householder loop (dlarfg convention, scaled norms
// and dot products), leaving V and R in place and its T at Ts[s]; the R's
// stacked (S nb x nb) are the next level's matrix.  The explicit Q is then
// built top down: slab s's rows are [X_s; 0] - V_s T_s V_s1^T X_s, X_s the
// s-th nb x nb block of the level above's Q (the identity at the top).
// ---------------------------------------------------------------------------
constexpr int TSQR_RP = 256;  // slab rows at most

__device__ __forceinline__ int tsqr_row(int s, int S, int M) {
  return (int)(((long long)s * M) / S);
}

// Householder QR (dlarfg convention) of the slab in LDS: nloc x nb, column
// stride RS, nb <= nloc <= PNT.  V below the diagonal and R on and above it
// in place, tau[j].  Per column the 256 threads split as (column k = t % 32,
// 32-row chunk t / 32): each sums its chunk of x.x (k == j) or x.a_k (k > j),
// the chunk partials are added in a fixed order; when max|x| leaves
// [1e-150, 1e150] or a sum is not finite the pass is repeated on x times a
// power of two near 1 / max|x| (dnrm2's purpose).  Then one row per thread
// for the rank-1 update.
constexpr int TSQR_TAIL = 8 * 34 + 4 * PNB + 16 + 2 * PNB * LDM + PNB * (PNB + 1) + 8;

__device__ void tsqr_slab_factor(double *slab, int RS, int nloc, int nb, double *tail) {
  double *red = tail;             // 8 chunks x 34 (32 sums, the chunk max)
  double *dots = red + 8 * 34;    // combined sums
  double *wv = dots + PNB;        // w_k
  double *tau = wv + PNB;         // tau_j
  double *cmax = tau + PNB;       // per-column flag space (unused)
  double *scal = cmax + PNB;      // [0] beta [1] tau [2] scale [3] dscale [4] max [5] sc [7] bad
  const int tid = threadIdx.x, kc = tid & 31, ch = tid >> 5;
  const int rb = ch * 32, re = min(nloc, rb + 32);
  (void)cmax;
  for (int j = 0; j < nb; j++) {
    for (int pass = 0; pass < 2; pass++) {
      double inv = 1.0;
      if (pass == 1) {
        const double mx = scal[4];
        if (!(mx != 0.0 && (mx < 1e-150 || mx > 1e150 || scal[7] != 0.0))) break;
        inv = ldexp(1.0, -ilogb(mx));
      }
      double acc = 0.0, mx = 0.0;
      if (kc >= j && kc < nb) {
        // fixed trip count, predicated: the 64 LDS reads issue back to back
        const int rlo = max(rb, j + 1);
#pragma unroll
        for (int i = 0; i < 32; i++) {
          const int r = rb + i;
          const bool ok = r >= rlo && r < re;
          const double x = ok ? slab[r + j * RS] * inv : 0.0;
          const double y = ok ? slab[r + kc * RS] * (kc == j ? inv : 1.0) : 0.0;
          acc += x * y;
          mx = fmax(mx, fabs(x));
        }
      }
      red[ch * 34 + kc] = acc;
      if (kc == j) red[ch * 34 + 32] = mx;
      __syncthreads();
      if (tid < PNB) {
        double sm = 0.0;
#pragma unroll
        for (int c = 0; c < 8; c++) sm += red[c * 34 + tid];
        dots[tid] = sm;
      }
      if (tid == PNB) {
        double m8 = 0.0;
#pragma unroll
        for (int c = 0; c < 8; c++) m8 = fmax(m8, red[c * 34 + 32]);
        if (pass == 0) scal[4] = m8;
        scal[5] = (pass == 0) ? 1.0 : ldexp(1.0, ilogb(scal[4]));
      }
      __syncthreads();
      if (pass == 0 && tid == 0) {
        int bad = 0;
        for (int k = j; k < nb; k++) bad |= !(fabs(dots[k]) <= 1.7976931348623157e308);
        scal[7] = bad ? 1.0 : 0.0;
      }
      __syncthreads();
    }
    if (tid == 0) {
      const double sc = scal[5];
      const double alpha = slab[j + j * RS], xnorm = sc * sqrt(dots[j]);
      double beta = alpha, t = 0.0, scale = 0.0, dscale = 0.0;
      if (xnorm != 0.0) {
        const double aa = fabs(alpha);
        const double w = fmax(aa, xnorm), z = fmin(aa, xnorm);
        const double nrm = (z == 0.0) ? w : w * sqrt(1.0 + (z / w) * (z / w));
        beta = -copysign(nrm, alpha);
        t = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
        dscale = sc / (alpha - beta);  // the dots are of x / sc
      }
      scal[0] = beta;
      scal[1] = t;
      scal[2] = scale;
      scal[3] = dscale;
      tau[j] = t;
    }
    __syncthreads();
    if (tid > j && tid < nb) wv[tid] = scal[1] * (slab[j + tid * RS] + dots[tid] * scal[3]);
    __syncthreads();
    const double t = scal[1];
    if (t != 0.0 && tid < nloc && tid >= j) {
      const int r = tid;
      double v = 1.0;
      if (r > j) {
        v = slab[r + j * RS] * scal[2];
        slab[r + j * RS] = v;
      } else {
        slab[j + j * RS] = scal[0];
      }
#pragma unroll
      for (int k = 0; k < PNB; k++)
        if (k > j && k < nb) slab[r + k * RS] -= v * wv[k];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(PNT) void tsqr_leaf_kernel(int M, int nb, double *X, int ldx, int S, int rs,
                                                        double *Ts, double *Y, int ldy) {
  extern __shared__ double smem[];
  const int s = blockIdx.x, tid = threadIdx.x;
  const int r0 = tsqr_row(s, S, M), nloc = tsqr_row(s + 1, S, M) - r0;
  double *slab = smem;
  double *tail = smem + slab_doubles(nb, rs);
  double *Gm = tail + 8 * 34 + 4 * PNB + 16;  // Gram of V (ld LDM)
  double *Rs = Gm + PNB * LDM;                 // R (ld LDM)
  double *Tl = Rs + PNB * LDM;                 // T (ld PNB + 1)
  const double *taus = tail + 8 * 34 + 2 * PNB;
  double *Xs = X + r0;
  // one row per thread (nloc <= PNT): all 32 loads in flight before the
  // first LDS write
  if (tid < nloc) {
    double row[PNB];
#pragma unroll
    for (int c = 0; c < PNB; c++) row[c] = (c < nb) ? Xs[tid + (size_t)c * ldx] : 0.0;
#pragma unroll
    for (int c = 0; c < PNB; c++)
      if (c < nb) slab[tid + c * rs] = row[c];
  }
  __syncthreads();
  tsqr_slab_factor(slab, rs, nloc, nb, tail);
  // R aside; the top block becomes V1 (unit lower) for the Gram V^T V
  for (int t = tid; t < PNB * PNB; t += PNT) {
    const int r = t % PNB, c = t / PNB;
    if (r < nb && c < nb) {
      const double x = slab[r + c * rs];
      Rs[r * LDM + c] = (r <= c) ? x : 0.0;
      slab[r + c * rs] = (r > c) ? x : (r == c ? 1.0 : 0.0);
    }
  }
  __syncthreads();
  slab_gram(slab, rs, nloc, nb, Gm);
  __syncthreads();
  // T = dlarft('F', 'C'): T(0:j, j) = -tau_j T(0:j, 0:j) G(0:j, j)
  for (int t = tid; t < PNB * (PNB + 1); t += PNT) Tl[t] = 0.0;
  __syncthreads();
  for (int j = 0; j < nb; j++) {
    const double tj = taus[j];
    double v = 0.0;
    if (tid < j && tj != 0.0)
      for (int q = tid; q < j; q++) v += Tl[tid * (PNB + 1) + q] * Gm[q * LDM + j];
    __syncthreads();
    if (tid < j) Tl[tid * (PNB + 1) + j] = -tj * v;
    if (tid == j) Tl[j * (PNB + 1) + j] = tj;
    __syncthreads();
  }
  // V and R back to global, T to Ts[s], R to Y
  if (tid < nloc) {
    const int lr = tid;
#pragma unroll
    for (int c = 0; c < PNB; c++)
      if (c < nb) Xs[lr + (size_t)c * ldx] = (lr < nb && lr <= c) ? Rs[lr * LDM + c] : slab[lr + c * rs];
  }
  double *Tg = Ts + (size_t)s * PNB * PNB;
  for (int t = tid; t < nb * nb; t += PNT) {
    const int r = t % nb, c = t / nb;
    Tg[r + c * PNB] = (r <= c) ? Tl[r * (PNB + 1) + c] : 0.0;
    Y[(size_t)s * nb + r + (size_t)c * ldy] = Rs[r * LDM + c];
  }
}

// Q rows of slab s of the level's M x nb matrix V (factored in place by the
// leaf kernel): [X_s; 0] - V_s (T_s (V_s1^T X_s)), X_s rows [s nb, s nb + nb)
// of Xq (the identity when Xq is null), into Qo rows [r0, r1)
__global__ __launch_bounds__(PNT) void tsqr_expand_kernel(int M, int nb, int S, const double *V, int ldv,
                                                          const double *Ts, const double *Xq, int ldx,
                                                          double *Qo, int ldq) {
  __shared__ double X[PNB][PNB + 1], W[PNB][PNB + 1], Z[PNB][PNB + 1], Tl[PNB][PNB + 1];
  const int s = blockIdx.x, tid = threadIdx.x;
  const int r0 = tsqr_row(s, S, M), nloc = tsqr_row(s + 1, S, M) - r0;
  const double *Vs = V + r0;
  const double *Tg = Ts + (size_t)s * PNB * PNB;
  for (int t = tid; t < PNB * PNB; t += PNT) {
    const int r = t % PNB, c = t / PNB;
    double x = 0.0, v = 0.0, tv = 0.0;
    if (r < nb && c < nb) {
      x = Xq ? Xq[(size_t)s * nb + r + (size_t)c * ldx] : (r == c ? 1.0 : 0.0);
      v = (r > c) ? Vs[r + (size_t)c * ldv] : (r == c ? 1.0 : 0.0);  // V_s1 (unit lower)
      tv = (r <= c) ? Tg[r + c * PNB] : 0.0;
    }
    X[r][c] = x;
    Z[r][c] = v;
    Tl[r][c] = tv;
  }
  __syncthreads();
  // W = V_s1^T X
  for (int t = tid; t < PNB * PNB; t += PNT) {
    const int p = t % PNB, c = t / PNB;
    double acc = 0.0;
    for (int r = p; r < nb; r++) acc += Z[r][p] * X[r][c];
    W[p][c] = acc;
  }
  __syncthreads();
  // Z = T W (T upper)
  for (int t = tid; t < PNB * PNB; t += PNT) {
    const int p = t % PNB, c = t / PNB;
    double acc = 0.0;
    for (int q = p; q < nb; q++) acc += Tl[p][q] * W[q][c];
    Z[p][c] = acc;
  }
  __syncthreads();
  // rows: [X; 0] - V Z, one row per thread, V's row read across its columns
  for (int r = tid; r < nloc; r += PNT) {
    double vr[PNB];
#pragma unroll
    for (int p = 0; p < PNB; p++)
      vr[p] = (p < nb && p <= r) ? (p == r ? 1.0 : Vs[r + (size_t)p * ldv]) : 0.0;
    for (int c = 0; c < nb; c++) {
      double acc = (r < nb) ? X[r][c] : 0.0;
#pragma unroll
      for (int p = 0; p < PNB; p++) acc -= vr[p] * Z[p][c];
      Qo[(size_t)(r0 + r) + (size_t)c * ldq] = acc;
    }
  }
}

__global__ void clear_flag_kernel(int *f) { *f = 0; }

size_t slab_doubles_h(int nb, int rs) { return (((size_t)nb * rs + 1) & ~(size_t)1); }
static_assert(STREAM_TAIL > HH_TAIL, "the streamed panel's fallback uses the LDS base as its tail");

size_t panel_smem_bytes(int nb, int rs) {
  return sizeof(double) * ((((size_t)nb * rs + 1) & ~(size_t)1) + LDS_TAIL);
}

}  // namespace

// persistent scratch size (doubles) the runtime allocates and zeroes once
extern "C" size_t ld_panel_scratch_doubles(void) { return 3 * (size_t)SCRATCH_ALL; }

// mode: 0 column-by-column Householder, 1 Cholesky-QR2 + reconstruction
// (with the column-by-column fallback inside)
// slabs of a panel launch: G workgroups of rp rows; taller than 256 LDS
// slabs (rp > rmax) the panel stays in global memory
static void panel_geom(int mode, int ml, int nb, int *G_, int *rp_, int *rmax_) {
  int rmax = LDS_SLAB_DOUBLES / nb;
  if (!(rmax & 1)) rmax--;
  int target = (mode == 1) ? ld_tune()->panel_rows_cqr : ld_tune()->panel_rows_per_wg;
  if (target < nb) target = nb;
  if (target > rmax) target = rmax;
  int G = (ml + target - 1) / target;
  if (G > 256) G = 256;
  int rp = (ml + G - 1) / G;
  if (rp < nb) rp = nb;
  G = (ml + rp - 1) / rp;
  *G_ = G;
  *rp_ = rp;
  *rmax_ = rmax;
}

static void panel_launch(int mode, int ml, int nb, double *A, int lda, double *T, int ldt, double *Q,
                         int ldq, int decline) {
  if (nb <= 0 || ml <= 0) return;
  if (nb > PNB) {
    ld_fatal("panel width above 32 reached the fused panel kernel");
    return;
  }
  // slab rows per workgroup: as many as LDS holds next to the fixed part,
  // with an odd column stride (LDS bank spread)
  int G, rp, rmax;
  panel_geom(mode, ml, nb, &G, &rp, &rmax);
  // taller than 256 LDS slabs: the column-by-column kernel on the panel in
  // global memory, 256 workgroups
  const bool glob = rp > rmax;
  // streamed Cholesky-QR2 panel: sub-slabs of the LDS that the fixed part
  // leaves, odd column stride
  const bool stream = glob && mode == 1;
  int srs = (LDS_DYN_BYTES / 8 - STREAM_TAIL - 2) / nb;
  if (!(srs & 1)) srs--;
  const int rs = stream ? srs : glob ? lda : (rp | 1);  // odd LDS column stride
  // Co-residency of all G workgroups is what the in-kernel barrier needs.
  // It is checked once from the occupancy calculator (one workgroup per CU
  // at the largest LDS size, G <= CU count); a plain launch then keeps all
  // G resident because the stream runs nothing else meanwhile.  The
  // cooperative launch API gives the same guarantee at a higher launch cost
  // and is used when LILA_PANEL_COOP=1 or when the check fails.
  static int attr_set = 0, ncu = 0, coop = 0;
  const size_t maxb = LDS_DYN_BYTES;
  if (!attr_set) {
    // the dynamic LDS asked for plus each kernel's static LDS (a few bytes:
    // LDS_DYN_BYTES leaves 16) must fit the 160 KiB of a CU; a failure here
    // is reported and cleared, not left pending for the next runtime call
    const void *kern[5] = {(const void *)panel_hh_kernel<false>, (const void *)panel_hh_kernel<true>,
                           (const void *)panel_cqr_kernel, (const void *)panel_hh_global_kernel,
                           (const void *)panel_cqr_stream_kernel};
    for (int q = 0; q < 5; q++) {
      const hipError_t e = hipFuncSetAttribute(kern[q], hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)maxb);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        fprintf(stderr, "lila_hip: panel kernel %d: %zu bytes of dynamic LDS refused (%s)\n", q,
                maxb, hipGetErrorString(e));
        ld_set_info(-5);
      }
    }
    int dev = 0, per_cu = 0, per_cu2 = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)panel_hh_kernel<false>, PNT,
                                                 maxb);
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu2, (const void *)panel_cqr_kernel, PNT, maxb);
    const char *e = getenv("LILA_PANEL_COOP");
    coop = (e && atoi(e) == 1) || per_cu < 1 || per_cu2 < 1 || ncu < 1;
    attr_set = 1;
  }
  if (ncu > 0 && G > ncu) {
    ld_fatal("panel needs more workgroups than compute units");
    return;
  }
  PanelArgs pa;
  pa.ml = ml; pa.nb = nb; pa.A = A; pa.lda = lda; pa.T = T; pa.ldt = ldt;
  pa.Q = Q; pa.ldq = ldq; pa.G = G; pa.rp = rp; pa.rs = rs;
  pa.decline = decline; pa.batch = 0;
  pa.sr = stream ? srs : rp;
  pa.scr = ld_panel_scratch();
  int *flags = ld_device_flags();
  pa.cnt = (unsigned *)(flags + 8);
  pa.bar = (unsigned *)(pa.scr + OFF_BAR);
  pa.err = flags;
  const void *fn;
  const char *cls;
  if (stream) {
    fn = (const void *)panel_cqr_stream_kernel;
    cls = "panel_cqr_stream_kernel";
  } else if (glob) {
    fn = (const void *)panel_hh_global_kernel;
    cls = "panel_hh_global_kernel";
  } else if (G == 1) {
    fn = (const void *)panel_hh_kernel<true>;
    cls = "panel_hh_kernel<true>";
  } else if (mode == 1) {
    fn = (const void *)panel_cqr_kernel;
    cls = "panel_cqr_kernel";
  } else {
    fn = (const void *)panel_hh_kernel<false>;
    cls = "panel_hh_kernel<false>";
  }
  void *args[] = {&pa};
  const size_t shb = stream ? sizeof(double) * (slab_doubles_h(nb, rs) + STREAM_TAIL)
                   : glob ? sizeof(double) * LDS_TAIL : panel_smem_bytes(nb, rs);
  const int ph = ld_prof_begin(cls);
  hipError_t e;
  // with two lanes the other lane's GEMMs may hold CUs while the panel
  // launches, so co-residency is asked of the cooperative launch
  if ((coop || ld_lanes() > 1) && G > 1) {
    e = hipLaunchCooperativeKernel(fn, dim3(G), dim3(PNT), args, (unsigned)shb, (hipStream_t)ld_stream());
  } else {
    e = hipLaunchKernel(fn, dim3(G), dim3(PNT), args, shb, (hipStream_t)ld_stream());
  }
  // algorithmic traffic: read the panel, write V/R (+ Q), write T
  ld_prof_end(ph, 2.0 * ml * nb * nb + (Q ? 2.0 * ml * nb * nb : 0.0),
              8.0 * ((Q ? 3.0 : 2.0) * ml * nb + (double)nb * nb));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fprintf(stderr, "lila_hip: panel launch failed (G=%d rp=%d nb=%d): %s\n", G, rp, nb,
            hipGetErrorString(e));
    ld_set_info(-1);
  }
}

extern "C" void ld_panel_qr(int ml, int nb, double *A, int lda, double *T, int ldt, double *Q,
                            int ldq) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  panel_launch(ld_tune()->panel_mode, ml, nb, A, lda, T, ldt, Q, ldq, 0);
}

// 64-wide Cholesky-QR2 panel: workgroups, rows per workgroup (at least 64,
// so that workgroup 0 holds the panel's top 64 rows) and the odd LDS column
// stride of the slab plus the 64 top rows; false when the panel is too tall
// for 256 such slabs
static bool panel64_geom(int ml, int *G_, int *rp_, int *rs_) {
  int rsmax = (LDS_DYN_BYTES / 8 - CQR64_TAIL - 2) / 64;
  if (!(rsmax & 1)) rsmax--;
  int rpmax = rsmax - 64;
  if (rpmax > PNT) rpmax = PNT;  // the fallback moves one slab row per thread
  if (rpmax < 64) return false;
  int target = ld_tune()->panel_rows_cqr;
  if (target < 64) target = 64;
  if (target > rpmax) target = rpmax;
  int G = (ml + target - 1) / target;
  if (G > 256) G = 256;
  int rp = (ml + G - 1) / G;
  if (rp < 64) rp = 64;
  if (rp > rpmax) return false;
  G = (ml + rp - 1) / rp;
  *G_ = G;
  *rp_ = rp;
  *rs_ = (rp + 64) | 1;
  return true;
}

// the ml x nb panel (32 < nb <= 64, no explicit Q) in one Cholesky-QR2
// launch; -1 (nothing launched) when it does not take the panel: panel mode
// 0, ml < 64, or taller than its LDS slabs
extern "C" int ld_panel_qr64(int ml, int nb, double *A, int lda, double *T, int ldt) {
  if (ld_failed()) return 0;  // an earlier allocation of this call failed
  if (nb <= PNB || nb > 64 || ml < 64 || ld_tune()->panel_mode != 1) return -1;
  int G, rp, rs;
  if (!panel64_geom(ml, &G, &rp, &rs)) return -1;
  static int attr = 0, coop = 0, ncu = 0;
  if (!attr) {
    const hipError_t e = hipFuncSetAttribute((const void *)panel_cqr64_kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_DYN_BYTES);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fprintf(stderr, "lila_hip: the 64-wide panel kernel cannot take %d bytes of LDS (%s); "
              "64-wide panels take the 32-wide recursion\n", (int)LDS_DYN_BYTES, hipGetErrorString(e));
      attr = -1;
    } else {
      int dev = 0, per_cu = 0;
      hipGetDevice(&dev);
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)panel_cqr64_kernel, PNT,
                                                   LDS_DYN_BYTES);
      const char *ev = getenv("LILA_PANEL_COOP");
      coop = (ev && atoi(ev) == 1) || per_cu < 1 || ncu < 1;
      attr = 1;
    }
  }
  if (attr < 0 || (ncu > 0 && G > ncu)) return -1;
  // one workgroup more than the slabs need, with no rows of its own: the
  // last workgroup forms T (every workgroup holds U and L after the LU)
  // while the others form and write V2, so T leaves the critical path
  if (ncu <= 0 || G + 1 <= ncu) G++;
  PanelArgs pa;
  memset(&pa, 0, sizeof pa);
  pa.ml = ml; pa.nb = nb; pa.A = A; pa.lda = lda; pa.T = T; pa.ldt = ldt;
  pa.Q = nullptr; pa.ldq = 0; pa.G = G; pa.rp = rp; pa.rs = rs; pa.sr = rp;
  pa.decline = 0; pa.batch = 0;
  // alternate scratch sets (see clear_set64): this launch's and the next's
  static int par64 = 0;
  double *const scr0 = ld_panel_scratch();
  pa.scr = scr0 + (size_t)SCRATCH_ALL * (1 + par64);
  pa.scr_next = scr0 + (size_t)SCRATCH_ALL * (2 - par64);
  par64 ^= 1;
  int *flags = ld_device_flags();
  pa.cnt = (unsigned *)(flags + 8);
  pa.bar = (unsigned *)(pa.scr + OFF_BAR);
  pa.err = flags;
  void *args[] = {&pa};
  const size_t shb = sizeof(double) * (slab_doubles_h(64, rs) + CQR64_TAIL);
  const int ph = ld_prof_begin("panel_cqr64_kernel");
  hipError_t e;
  if ((coop || ld_lanes() > 1) && G > 1)
    e = hipLaunchCooperativeKernel((const void *)panel_cqr64_kernel, dim3(G), dim3(PNT), args, (unsigned)shb,
                                   (hipStream_t)ld_stream());
  else
    e = hipLaunchKernel((const void *)panel_cqr64_kernel, dim3(G), dim3(PNT), args, shb,
                        (hipStream_t)ld_stream());
  ld_prof_end(ph, 2.0 * ml * nb * nb, 8.0 * (2.0 * ml * nb + (double)nb * nb));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fprintf(stderr, "lila_hip: 64-wide panel launch failed (G=%d rp=%d nb=%d): %s\n", G, rp, nb,
            hipGetErrorString(e));
    ld_set_info(-1);
  }
  return 0;
}

extern "C" void ld_panel_qr_exact(int ml, int nb, double *A, int lda, double *T, int ldt, double *Q,
                                  int ldq) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  panel_launch(0, ml, nb, A, lda, T, ldt, Q, ldq, 0);
}

// 1 when an ml x nb panel is taller than the workgroups' LDS slabs can hold
extern "C" int ld_panel_is_tall(int ml, int nb) {
  if (nb <= 0 || nb > PNB || ml <= 0) return 0;
  int G, rp, rmax;
  panel_geom(ld_tune()->panel_mode, ml, nb, &G, &rp, &rmax);
  return rp > rmax;
}

// a tall panel by the streamed Cholesky-QR2 kernel without its column-loop
// fallback: 0 when it factored the panel, 1 when a Cholesky or LU acceptance
// test failed and the panel was left untouched (waits for the launch)
extern "C" int ld_panel_qr_tall_cqr(int ml, int nb, double *A, int lda, double *T, int ldt, double *Q,
                                    int ldq) {
  if (ld_failed()) return 0;
  panel_launch(1, ml, nb, A, lda, T, ldt, Q, ldq, 1);
  int *flags = ld_device_flags();
  int h = 0;
  hipStream_t st = (hipStream_t)ld_stream();
  if (hipMemcpyAsync(&h, flags + 2, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    (void)hipGetLastError();
    ld_set_info(-1);
    return 0;
  }
  if (h) {
    hipLaunchKernelGGL(clear_flag_kernel, dim3(1), dim3(1), 0, st, flags + 2);
    return 1;
  }
  return 0;
}

// TSQR levels (slab count S of an M x nb level; M >= S nb)
extern "C" int ld_tsqr_slabs(int M, int nb) {
  static int rp = 0;
  if (!rp) {
    const char *e = getenv("LILA_TSQR_RP");
    rp = e ? atoi(e) : TSQR_RP;
    if (rp < 64 || rp > TSQR_RP) rp = TSQR_RP;
  }
  int S = (M + rp - 1) / rp;
  if (S < 1) S = 1;
  while (S > 1 && M / S < nb) S--;
  return S;
}

extern "C" void ld_tsqr_leaf(int M, int nb, double *X, int ldx, int S, double *Ts, double *Y,
                             int ldy) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (M <= 0 || nb <= 0 || nb > PNB) return;
  const int maxrows = (M + S - 1) / S;
  if (maxrows > PNT || M / S < nb) {
    ld_fatal("tsqr: slab outside [nb, 256] rows");
    return;
  }
  const int rs = maxrows | 1;
  static int attr = 0;
  const size_t shb = sizeof(double) * (slab_doubles_h(nb, rs) + TSQR_TAIL);
  if (!attr) {
    if (hipFuncSetAttribute((const void *)tsqr_leaf_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)LDS_DYN_BYTES) != hipSuccess) {
      (void)hipGetLastError();
      ld_fatal("tsqr: dynamic LDS refused");
      return;
    }
    attr = 1;
  }
  const int ph = ld_prof_begin("tsqr_leaf_kernel");
  hipLaunchKernelGGL(tsqr_leaf_kernel, dim3(S), dim3(PNT), shb, (hipStream_t)ld_stream(), M, nb, X, ldx,
                     S, rs, Ts, Y, ldy);
  ld_prof_end(ph, 2.0 * M * nb * nb, 8.0 * (2.0 * M * nb + 2.0 * S * nb * nb));
}

extern "C" void ld_tsqr_expand(int M, int nb, int S, const double *V, int ldv, const double *Ts,
                               const double *Xq, int ldx, double *Qo, int ldq) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (M <= 0 || nb <= 0) return;
  const int ph = ld_prof_begin("tsqr_expand_kernel");
  hipLaunchKernelGGL(tsqr_expand_kernel, dim3(S), dim3(PNT), 0, (hipStream_t)ld_stream(), M, nb, S, V,
                     ldv, Ts, Xq, ldx, Qo, ldq);
  ld_prof_end(ph, 2.0 * M * nb * nb, 8.0 * (2.0 * M * nb));
}
