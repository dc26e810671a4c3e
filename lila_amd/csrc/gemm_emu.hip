// gemm_emu.hip -- FP64 products C = alpha op(A) B + beta C emulated on int8
// MFMA (v_mfma_i32_32x32x32_i8) by the Chinese-remainder scheme of Ozaki,
// Uchino and Imamura ("Ozaki scheme II", 2024/25):
//
//   1. every row i of op(A) is scaled by 2^sa_i and every column j of B by
//      2^sb_j so that its largest entry has magnitude in [2^(kb-1), 2^kb),
//      and rounded to an integer: A', B' with |A'|, |B'| <= 2^kb;
//   2. for pairwise coprime moduli p_1 .. p_N the symmetric residues
//      A' mod p_t and B' mod p_t (all in [-128, 127]) are multiplied on int8
//      MFMA with exact int32 accumulation, C_t = A'B' mod p_t;
//   3. X = A'B' is the integer of |X| < P/2 (P = p_1 ... p_N) with those
//      residues: X = sum_t c_t q_t mod P, q_t = (P/p_t) ((P/p_t)^-1 mod p_t),
//      c_t = C_t reduced to |c_t| < 1.5 p_t.  q_t is split as q_t1 + q_t2
//      with q_t1 a multiple of 2^e of few enough bits that sum_t c_t q_t1 is
//      exact in FP64; with r = round(sum / P) and P = P1 + P2 split the same
//      way, X = (S1 - r P1) + (S2 - r P2), the first difference exact;
//   4. C = alpha 2^-(sa_i + sb_j) X + beta C.
//
// N is the smallest count with K 2^(2 kb) < 0.42 P, so step 3 recovers A'B'
// exactly up to the FP64 rounding of S2 (absolute ~2^41 against |X| up to
// ~2^(2 kb + log2 K)); kb = 53 and K <= 1024 take 15 moduli.  The only other rounding is step 1's (relative 2^-kb
// of the row's / column's largest entry): with kb = 53 the product is as
// accurate as an FP64 one in the norm-wise sense.
//
// Data: for each modulus the residues of op(A) are stored row by row (k
// contiguous, Kp bytes per row, Kp a multiple of 128, rows padded to a
// multiple of 128 with zeros), those of B column by column the same way.
// The GEMM kernel keeps a 128 x 128 output tile in registers for all moduli:
// 8 waves (2 per SIMD) of 32 x 64, int32 accumulators reset per modulus and
// folded into two FP64 CRT sums per entry at the modulus's last k step.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "lila_dev.h"

namespace {

// pairwise coprime moduli <= 256 (symmetric residues fit int8); P of all
// 16 is ~2^126.4, below the 2^127 the host's 128-bit CRT set-up allows
constexpr int EMU_MAXMOD = 16;
constexpr int EMU_MODS[EMU_MAXMOD] = {256, 255, 253, 251, 247, 241, 239, 233,
                                      229, 227, 223, 217, 211, 199, 197, 193};

struct EmuArgs {
  int M, N, Kp, nmod, ksteps, zero_rows;
  double alpha, beta;
  const int8_t *Ap;
  const int8_t *Bp;
  size_t sA, sB;  // bytes per modulus
  const int *sa;
  const int *sb;
  double *C;
  int ldc;
  double p[EMU_MAXMOD], invp[EMU_MAXMOD], q1[EMU_MAXMOD], q2[EMU_MAXMOD];
  double P1, P2, invP;
};

struct EmuMods {
  int nmod;
  double p[EMU_MAXMOD], invp[EMU_MAXMOD], lo[EMU_MAXMOD], hi[EMU_MAXMOD];
};

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_void;

constexpr int EM_BM = 128, EM_BN = 128, EM_BK = 128;  // EM_BK: k bytes per stage
constexpr int EM_NT = 512;
constexpr int EM_ASTAGE = EM_BM * EM_BK;  // bytes
constexpr int EM_STAGE = EM_ASTAGE + EM_BN * EM_BK;
constexpr int EM_NBUF = 4;                             // stages: 3 in flight while one is read
constexpr size_t EM_LDS = EM_NBUF * (size_t)EM_STAGE;  // 128 KiB
constexpr int EM_LOADS = 4;                            // load instructions per thread per stage
// s_waitcnt immediate (gfx9 encoding): at most n vector-memory operations
// of this wave outstanding and none of its LDS operations (expcnt not waited)
constexpr int em_wait(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (0 << 8); }

// the workgroup barrier of the stage pipeline: this wave's loads of the
// stage about to be read are complete (at most n later loads outstanding)
// and its LDS reads of the previous stage are done, then s_barrier.  Unlike
// __syncthreads() (whose fence waits for every outstanding load, the
// prefetched stages included) only the compiler is fenced.
__device__ __forceinline__ void em_stage_barrier(int n) {
  if (n >= 8) __builtin_amdgcn_s_waitcnt(em_wait(8));
  else if (n >= 4) __builtin_amdgcn_s_waitcnt(em_wait(4));
  else __builtin_amdgcn_s_waitcnt(em_wait(0));
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// LDS slot of 16-byte chunk c of row r of a stage operand: rows of 128 bytes,
// the chunk index XOR-ed with bits 1..3 of the row so that the 32 rows one
// MFMA operand read touches spread over the banks
__device__ __forceinline__ int em_slot(int r, int c) { return r * 8 + (c ^ ((r >> 1) & 7)); }

__device__ __forceinline__ void em_load(const int8_t *gp, char *lp) {
  __builtin_amdgcn_global_load_lds((const void *)gp, (lds_void *)lp, 16, 0, 0);
}

__global__ __launch_bounds__(EM_NT, 1) void emu_gemm_kernel(EmuArgs a) {
  extern __shared__ __attribute__((aligned(16))) char es[];
  int tx = blockIdx.x, ty = blockIdx.y;
  {
    // XCD-aware tile order: workgroups are dealt round-robin over the 8
    // XCDs; give each XCD a contiguous eighth of the tiles in 8 x 8 blocks
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
  const int wm = wave & 3, wn = wave >> 2;  // wave tile: rows 32 wm, columns 64 wn
  const int m0 = tx * EM_BM, n0 = ty * EM_BN;
  const int total = a.nmod * a.ksteps;

  // one stage: 1024 16-byte chunks of A and of B, two of each per thread;
  // load j of wave w fills slots j * 512 + w * 64 + lane
  auto issue = [&](int buf, int step) {
    const int t = step / a.ksteps, k0 = (step - t * a.ksteps) * EM_BK;
    char *as = es + buf * EM_STAGE, *bs = as + EM_ASTAGE;
    const int8_t *Ag = a.Ap + (size_t)t * a.sA + (size_t)m0 * a.Kp + k0;
    const int8_t *Bg = a.Bp + (size_t)t * a.sB + (size_t)n0 * a.Kp + k0;
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int s = j * 512 + tid, r = s >> 3, c = (s & 7) ^ ((r >> 1) & 7);
      const int base = (j * 512 + wave * 64) * 16;
      em_load(Ag + (size_t)r * a.Kp + 16 * c, as + base);
      em_load(Bg + (size_t)r * a.Kp + 16 * c, bs + base);
    }
  };

  double S1[32], S2[32];
#pragma unroll
  for (int e = 0; e < 32; e++) { S1[e] = 0.0; S2[e] = 0.0; }
  v16i acc[2];
  const int arow = wm * 32 + (lane & 31);
  const int brow0 = wn * 64 + (lane & 31), brow1 = brow0 + 32;
  const int khalf = lane >> 5;

  // stages 0 .. NBUF-2 in flight before the loop; at step s the wave waits
  // for its loads of stage s (those of the NBUF-2 later stages may still be
  // outstanding), the barrier makes every wave's stage-s data visible and
  // frees the buffer stage s-1 used, which then takes stage s + NBUF - 1
  for (int s0 = 0; s0 < EM_NBUF - 1; s0++)
    if (s0 < total) issue(s0, s0);
  for (int step = 0; step < total; step++) {
    const int t = step / a.ksteps, ks = step - t * a.ksteps;
    const int ahead = min(total - 1 - step, EM_NBUF - 2);  // later stages already issued
    em_stage_barrier(ahead * EM_LOADS);
    if (step + EM_NBUF - 1 < total) issue((step + EM_NBUF - 1) % EM_NBUF, step + EM_NBUF - 1);
    if (ks == 0) {
#pragma unroll
      for (int e = 0; e < 16; e++) { acc[0][e] = 0; acc[1][e] = 0; }
    }
    const char *as = es + (step % EM_NBUF) * EM_STAGE, *bs = as + EM_ASTAGE;
#pragma unroll
    for (int sub = 0; sub < EM_BK / 32; sub++) {
      const int c = 2 * sub + khalf;
      const v4i af = *(const v4i *)(as + 16 * em_slot(arow, c));
      const v4i b0 = *(const v4i *)(bs + 16 * em_slot(brow0, c));
      const v4i b1 = *(const v4i *)(bs + 16 * em_slot(brow1, c));
      acc[0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(b0, af, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(b1, af, acc[1], 0, 0, 0);
    }
    if (ks == a.ksteps - 1) {
      // fold modulus t into the CRT sums: c = C_t - p round(C_t / p), exact,
      // |c| < 1.5 p (the quotient may be off by one)
      const double p = a.p[t], ip = a.invp[t], q1 = a.q1[t], q2 = a.q2[t];
#pragma unroll
      for (int nb = 0; nb < 2; nb++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
          const double x = (double)acc[nb][v];
          const double q = rint(x * ip);
          const double c = fma(-q, p, x);
          S1[nb * 16 + v] = fma(c, q1, S1[nb * 16 + v]);
          S2[nb * 16 + v] = fma(c, q2, S2[nb * 16 + v]);
        }
    }
  }

  // X = (S1 - r P1) + (S2 - r P2), then C = alpha 2^-(sa + sb) X + beta C.
  // Lane l, register v of n-block nb: row m = l % 32 of the wave's rows,
  // column 32 nb + 8 (v / 4) + 4 (l / 32) + v % 4 of its columns.
  const int m = m0 + arow;
  if (m < a.M) {
    const int sam = a.sa[m];
#pragma unroll
    for (int nb = 0; nb < 2; nb++)
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int n = n0 + wn * 64 + nb * 32 + 8 * (v >> 2) + 4 * khalf + (v & 3);
        if (n < a.N) {
          const double s1 = S1[nb * 16 + v], s2 = S2[nb * 16 + v];
          const double r = rint((s1 + s2) * a.invP);
          const double X = fma(-r, a.P1, s1) + fma(-r, a.P2, s2);
          double o = a.alpha * ldexp(X, -(sam + a.sb[n]));
          double *cp = a.C + (size_t)m + (size_t)n * a.ldc;
          if (a.beta != 0.0 && m >= a.zero_rows) o = fma(a.beta, *cp, o);
          *cp = o;
        }
      }
  }
}

// scale exponent of each contiguous vector (a column of a column-major
// matrix, length K at X + v ldx): the largest entry times 2^s lies in
// [2^(kb-1), 2^kb); 0 for a zero vector.  One wave per vector.
__global__ void emu_scale_vec_kernel(int nvec, int K, const double *__restrict__ X, int ldx, int kb,
                                     int *__restrict__ s) {
  const int v = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (v >= nvec) return;
  const double *x = X + (size_t)v * ldx;
  double mx = 0.0;
  for (int k = lane; k < K; k += 64) mx = fmax(mx, fabs(x[k]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  if (lane == 0) s[v] = (mx > 0.0) ? kb - 1 - ilogb(mx) : 0;
}

// the same for the rows of a column-major M x K matrix (row i strided by
// lda), in two launches: partial maxima of 256 rows x 64 columns per
// workgroup (a thread per row, loads coalesced over the rows) combined by an
// atomic max on the bit patterns (non-negative doubles order like their
// bits) into e[] cleared beforehand, then the exponents
__global__ void emu_rowmax_kernel(int M, int K, const double *__restrict__ A, int lda,
                                  unsigned long long *__restrict__ e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const int k0 = blockIdx.y * 64, k1 = min(K, k0 + 64);
  double mx = 0.0;
  for (int k = k0; k < k1; k++) mx = fmax(mx, fabs(A[(size_t)i + (size_t)k * lda]));
  if (mx > 0.0) atomicMax(&e[i], (unsigned long long)__double_as_longlong(mx));
}
__global__ void emu_rowscale_kernel(int M, const unsigned long long *__restrict__ e, int kb,
                                    int *__restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const double mx = __longlong_as_double((long long)e[i]);
  s[i] = (mx > 0.0) ? kb - 1 - ilogb(mx) : 0;
}
__global__ void emu_clear_kernel(long n, unsigned long long *__restrict__ e) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) e[i] = 0ull;
}

// symmetric residues of 16 integers (as doubles, |x| <= 2^53) modulo each
// modulus, 16 bytes per modulus to P + t stride + off
__device__ __forceinline__ void emu_put16(const double (&x)[16], const EmuMods &md, int8_t *P,
                                          size_t stride, size_t off) {
  for (int t = 0; t < md.nmod; t++) {
    const double p = md.p[t], ip = md.invp[t], lo = md.lo[t], hi = md.hi[t];
    int w[4] = {0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < 16; e++) {
      // x - p round(x / p) is exact; the quotient can be off by one, so the
      // result is brought into [lo, hi] with one more step either way
      double r = fma(-rint(x[e] * ip), p, x[e]);
      r = (r > hi) ? r - p : ((r < lo) ? r + p : r);
      w[e >> 2] |= ((int)r & 0xff) << (8 * (e & 3));
    }
    *(v4i *)(P + (size_t)t * stride + off) = v4i{w[0], w[1], w[2], w[3]};
  }
}

// residues of contiguous vectors (vector v at X + v ldx, K long, scale s[v])
// into P (vector v at v Kp, zero beyond K and for v >= nvec up to nvec_pad):
// one thread per 16-byte chunk
__global__ void emu_pack_vec_kernel(int nvec, int nvec_pad, int K, int Kp, const double *__restrict__ X,
                                    int ldx, const int *__restrict__ s, EmuMods md, int8_t *P,
                                    size_t stride) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int cpv = Kp / 16;
  if (idx >= (long)nvec_pad * cpv) return;
  const int v = (int)(idx / cpv), k0 = (int)(idx % cpv) * 16;
  double x[16];
  const bool live = v < nvec;
  const int sv = live ? s[v] : 0;
#pragma unroll
  for (int e = 0; e < 16; e++) {
    const int k = k0 + e;
    x[e] = (live && k < K) ? rint(ldexp(X[(size_t)v * ldx + k], sv)) : 0.0;
  }
  emu_put16(x, md, P, stride, (size_t)v * Kp + k0);
}

// residues of the rows of a column-major M x K matrix (row i strided by
// lda) into P (row i at i Kp): 64 x 64 tiles transposed through LDS
__global__ __launch_bounds__(256) void emu_pack_rows_kernel(int M, int K, int Kp, const double *__restrict__ A,
                                                            int lda, const int *__restrict__ s, EmuMods md,
                                                            int8_t *P, size_t stride) {
  __shared__ double tl[64][65];
  const int i0 = blockIdx.x * 64, k0 = blockIdx.y * 64, tid = threadIdx.x;
  {
    const int r = tid & 63;
    const int i = i0 + r;
    const int sv = (i < M) ? s[i] : 0;
#pragma unroll 4
    for (int kk = tid >> 6; kk < 64; kk += 4) {
      const int k = k0 + kk;
      tl[kk][r] = (i < M && k < K) ? rint(ldexp(A[(size_t)i + (size_t)k * lda], sv)) : 0.0;
    }
  }
  __syncthreads();
  const int r = tid >> 2, ch = tid & 3;
  double x[16];
#pragma unroll
  for (int e = 0; e < 16; e++) x[e] = tl[ch * 16 + e][r];
  emu_put16(x, md, P, stride, (size_t)(i0 + r) * Kp + k0 + ch * 16);
}

// ---- host side ------------------------------------------------------------

typedef unsigned __int128 u128;

struct EmuSetup {
  int nmod;
  double logP;
  double p[EMU_MAXMOD], invp[EMU_MAXMOD], q1[EMU_MAXMOD], q2[EMU_MAXMOD];
  double P1, P2, invP;
};

static long long inv_mod(long long a, long long m) {  // a^-1 mod m, gcd(a, m) = 1
  long long g = m, x = 0, x1 = 1, aa = a % m;
  if (aa < 0) aa += m;
  long long b = aa;
  while (b) {
    const long long q = g / b;
    long long t = g - q * b; g = b; b = t;
    t = x - q * x1; x = x1; x1 = t;
  }
  x %= m;
  return x < 0 ? x + m : x;
}

// high part of v: v rounded down to a multiple of 2^e, as a double (exact
// when v >> e has at most 53 bits), and the rest
static void split_u128(u128 v, int e, double &hi, double &lo) {
  const u128 h = (v >> e) << e;
  hi = ldexp((double)(unsigned long long)(v >> e), e);
  const u128 r = v - h;
  lo = (double)(unsigned long long)(r >> 64) * 18446744073709551616.0 + (double)(unsigned long long)r;
}

static int bits_u128(u128 v) {
  int b = 0;
  while (v) { v >>= 1; b++; }
  return b;
}

static const EmuSetup *emu_setup(int nmod) {
  static EmuSetup cache[EMU_MAXMOD + 1];
  static int done[EMU_MAXMOD + 1];
  if (nmod < 2 || nmod > EMU_MAXMOD) return nullptr;
  EmuSetup &s = cache[nmod];
  if (done[nmod]) return &s;
  u128 P = 1;
  for (int t = 0; t < nmod; t++) P *= (u128)EMU_MODS[t];
  s.nmod = nmod;
  s.logP = 0.0;
  for (int t = 0; t < nmod; t++) s.logP += log2((double)EMU_MODS[t]);
  // q_t1 keeps qb bits above 2^e: |c_t| < 1.5 p <= 384 < 2^9, so the nmod
  // products c_t q_t1 and their sum stay below 2^53
  int lg = 0;
  while ((1 << lg) < nmod) lg++;
  const int qb = 53 - 9 - lg - 1;
  const int e = bits_u128(P) - qb;
  for (int t = 0; t < nmod; t++) {
    const long long pt = EMU_MODS[t];
    const u128 Mt = P / (u128)pt;
    const long long y = inv_mod((long long)(Mt % (u128)pt), pt);
    const u128 q = (Mt * (u128)y) % P;
    s.p[t] = (double)pt;
    s.invp[t] = 1.0 / (double)pt;
    split_u128(q, e, s.q1[t], s.q2[t]);
  }
  split_u128(P, e, s.P1, s.P2);
  s.invP = 1.0 / ((double)(unsigned long long)(P >> 64) * 18446744073709551616.0 +
                  (double)(unsigned long long)P);
  done[nmod] = 1;
  return &s;
}

inline hipStream_t S() { return (hipStream_t)ld_stream(); }

}  // namespace

// the modulus count for a product over K with kb-bit operands: the smallest
// count with K 2^(2 kb) < 0.42 P, so that |X| / P <= 0.42 and round(S / P)
// is the right multiple of P
extern "C" int ld_emu_moduli(int K, int kb) {
  int lk = 0;
  while ((1L << lk) < (long)K) lk++;
  const double need = 2.0 * kb + lk + 1.25;
  for (int n = 2; n <= EMU_MAXMOD; n++)
    if (emu_setup(n)->logP >= need) return n;
  return -1;
}

// C = alpha op(A) B + beta C (rows < zero_rows take beta = 0) emulated on
// int8 MFMA; op(A) = A or A^T, B not transposed, general operands.  Returns
// 0 when done, -1 when the product is not one this path takes (the caller
// then runs the FP64 kernels).
extern "C" int ld_gemm_emu(const ld_gemm_t *g, int kb) {
  if (ld_failed()) return 0;
  const int M = g->m, N = g->n, K = g->k;
  if (M <= 0 || N <= 0) return 0;
  if (g->tb || g->sa != LD_GEN || g->sb != LD_GEN || g->c_upper || K <= 0 || K > 65536 ||
      g->alpha == 0.0)
    return -1;
  if (kb < 8 || kb > 53) kb = 53;
  const int nmod = ld_emu_moduli(K, kb);
  const EmuSetup *su = emu_setup(nmod);
  if (!su) return -1;
  static int attr = 0;
  if (!attr) {
    if (hipFuncSetAttribute((const void *)emu_gemm_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)EM_LDS) != hipSuccess) {
      (void)hipGetLastError();
      return -1;
    }
    attr = 1;
  }
  const int Mp = (M + EM_BM - 1) / EM_BM * EM_BM, Np = (N + EM_BN - 1) / EM_BN * EM_BN;
  const int Kp = (K + EM_BK - 1) / EM_BK * EM_BK;
  const size_t sA = (size_t)Mp * Kp, sB = (size_t)Np * Kp;
  const size_t mark = ld_ws_mark();
  int8_t *Ap = (int8_t *)ld_ws_alloc(sA * nmod);
  int8_t *Bp = (int8_t *)ld_ws_alloc(sB * nmod);
  int *sa = (int *)ld_ws_alloc(sizeof(int) * (size_t)(Mp + Np));
  unsigned long long *emax = g->ta ? nullptr : (unsigned long long *)ld_ws_alloc(sizeof(unsigned long long) * (size_t)M);
  if (ld_failed()) {
    ld_ws_release(mark);
    return 0;
  }
  int *sb = sa + Mp;
  EmuMods md;
  md.nmod = nmod;
  for (int t = 0; t < nmod; t++) {
    md.p[t] = su->p[t];
    md.invp[t] = su->invp[t];
    md.lo[t] = (EMU_MODS[t] == 256) ? -128.0 : -(double)((EMU_MODS[t] - 1) / 2);
    md.hi[t] = (EMU_MODS[t] == 256) ? 127.0 : (double)((EMU_MODS[t] - 1) / 2);
  }
  const int ph0 = ld_prof_begin("emu_pack");
  // op(A): rows of A^T are columns of A (contiguous); rows of A are strided
  if (g->ta) {
    hipLaunchKernelGGL(emu_scale_vec_kernel, dim3((M + 3) / 4), dim3(256), 0, S(), M, K, g->A, g->lda, kb, sa);
    const long nth = (long)Mp * (Kp / 16);
    hipLaunchKernelGGL(emu_pack_vec_kernel, dim3((unsigned)((nth + 255) / 256)), dim3(256), 0, S(), M, Mp, K, Kp,
                       g->A, g->lda, (const int *)sa, md, Ap, sA);
  } else {
    hipLaunchKernelGGL(emu_clear_kernel, dim3((M + 255) / 256), dim3(256), 0, S(), (long)M, emax);
    hipLaunchKernelGGL(emu_rowmax_kernel, dim3((M + 255) / 256, (K + 63) / 64), dim3(256), 0, S(), M, K, g->A,
                       g->lda, emax);
    hipLaunchKernelGGL(emu_rowscale_kernel, dim3((M + 255) / 256), dim3(256), 0, S(), M,
                       (const unsigned long long *)emax, kb, sa);
    hipLaunchKernelGGL(emu_pack_rows_kernel, dim3(Mp / 64, Kp / 64), dim3(256), 0, S(), M, K, Kp, g->A, g->lda,
                       (const int *)sa, md, Ap, sA);
  }
  hipLaunchKernelGGL(emu_scale_vec_kernel, dim3((N + 3) / 4), dim3(256), 0, S(), N, K, g->B, g->ldb, kb, sb);
  {
    const long nth = (long)Np * (Kp / 16);
    hipLaunchKernelGGL(emu_pack_vec_kernel, dim3((unsigned)((nth + 255) / 256)), dim3(256), 0, S(), N, Np, K, Kp,
                       g->B, g->ldb, (const int *)sb, md, Bp, sB);
  }
  ld_prof_end(ph0, 0.0, 8.0 * ((double)M * K + (double)K * N) + (double)nmod * (sA + sB));
  EmuArgs a;
  memset(&a, 0, sizeof a);
  a.M = M; a.N = N; a.Kp = Kp; a.nmod = nmod; a.ksteps = Kp / EM_BK;
  a.zero_rows = g->zero_rows;
  a.alpha = g->alpha; a.beta = g->beta;
  a.Ap = Ap; a.Bp = Bp; a.sA = sA; a.sB = sB; a.sa = sa; a.sb = sb;
  a.C = g->C; a.ldc = g->ldc;
  for (int t = 0; t < nmod; t++) {
    a.p[t] = su->p[t]; a.invp[t] = su->invp[t]; a.q1[t] = su->q1[t]; a.q2[t] = su->q2[t];
  }
  a.P1 = su->P1; a.P2 = su->P2; a.invP = su->invP;
  const int ph = ld_prof_begin("emu_gemm_kernel");
  hipLaunchKernelGGL(emu_gemm_kernel, dim3(Mp / EM_BM, Np / EM_BN), dim3(EM_NT), EM_LDS, S(), a);
  ld_prof_end(ph, 2.0 * M * N * (double)K,
              (double)nmod * (sA + sB) + 8.0 * (g->beta != 0.0 ? 2.0 : 1.0) * M * N);
  ld_ws_release(mark);
  return 0;
}
