// node_qt.hip -- the two block steps of a narrow dgeqr3 node, one launch
// each (gfx950).
//
// dgeqr3 (lila-c/dgeqr3.c:3-114) splits n columns into n1 = n/2 and
// n2 = n - n1, factors the left part, applies its reflectors to the right
// part, factors that, and joins the two T factors.  For n <= 128 both halves
// are at most 64 wide and the two block steps are thin products over all ml
// rows followed by 64 x 64 work:
//   update (lila-c/dgeqr3.c:41-70):  C <- C - V (T1^T (V^T C))
//          V: ml x n1 unit lower (the left half's reflectors), C: ml x n2
//   join   (lila-c/dgeqr3.c:78-110): T12 = -T1 (V1^T V2) T2
//          V1, V2: the ml2 rows below the left half's diagonal block
// As library GEMMs each step is three launches (a split-K product, the
// reduction fused with the triangular factor, the last product), and each
// launch boundary costs several microseconds around little arithmetic.  Here
// each step is ONE launch of G <= CUs workgroups (all resident):
//   1. workgroup g forms the partial product of its rows (64-row chunks
//      staged in LDS, v_mfma_f64_16x16x4f64) and stores it to its own slot;
//   2. after a grid barrier the n2 columns of the sum are dealt out to the
//      workgroups; each sums its columns over the G slots in a fixed order
//      (no atomics: the result does not vary from run to run) and multiplies
//      them by T1^T (update) or T1 (join);
//   3. update: after a second barrier every workgroup applies the 64 x 64
//      W to its rows (C -= V W, from the chunk still in LDS).  join: the last
//      workgroup out of step 2 (an arrival counter; the others exit) forms
//      T12 = -(T1 X) T2 on MFMA.
// (One 64 x 64 sum by FP64 atomics instead of the slots measured slower: a
// few hundred adds to each address are performed one after another.)
// The barrier counters alternate between two sets: a launch clears the set
// the next launch will use (the launch before it, which used that set, has
// finished on the stream).  Slots and the middle factor are overwritten
// whole by every launch.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "lila_dev.h"
#include "gridsync.h"

extern "C" int *ld_device_flags(void);
extern "C" double *ld_node_scratch(void);

typedef double d4_t __attribute__((ext_vector_type(4)));

namespace {
using namespace gridsync;

constexpr int NT = 256;      // threads per workgroup (four waves)
constexpr int NW = 64;       // widest half
constexpr int RC = 64;       // rows per LDS chunk
constexpr int LDR = RC + 1;  // column stride of a staged chunk (odd: no bank conflicts between columns)
constexpr int LDW = NW + 1;  // column stride of the 64 x 64 LDS matrices
constexpr int GMAX = 256;    // workgroups at most
// scratch layout (doubles): the partial slots, the 64 x 64 middle factor
// (column-major, ld 64), two sets of barrier counters (BAR_GROUPS + 1
// counters, then the join's arrival counter)
constexpr size_t OFF_PART = 0;                                // slot of (column j, workgroup g): 64 doubles at (j * G + g) * 64
constexpr size_t OFF_W = OFF_PART + (size_t)NW * GMAX * NW;
constexpr size_t OFF_BAR = OFF_W + NW * NW;
constexpr size_t BAR_WORDS = (size_t)(BAR_GROUPS + 2) * BAR_STRIDE;  // unsigned ints of a set
constexpr size_t NODE_SCRATCH = OFF_BAR + BAR_WORDS;                 // doubles (two sets of unsigned ints)
// LDS (doubles): two staged chunks, two 64 x 64 matrices, the reduction
// rows, the summed column and a flag; the join's tail takes a third 64 x 64
// matrix in the left chunk's place
constexpr int LDS_L = 0, LDS_R = LDS_L + NW * LDR, LDS_W = LDS_R + NW * LDR, LDS_T = LDS_W + NW * LDW,
              LDS_RED = LDS_T + NW * LDW, LDS_X = LDS_RED + 4 * NW, LDS_FLAG = LDS_X + NW,
              LDS_U = LDS_L, LDS_DOUBLES = LDS_FLAG + 2;
constexpr int LDS_BYTES = LDS_DOUBLES * 8;
static_assert(LDS_BYTES <= 160 * 1024, "LDS of one workgroup");
static_assert(NW * LDW <= NW * LDR, "the join's third matrix in the left chunk's place");

struct NodeArgs {
  int ml;                               // rows of both thin operands
  int n1, n2;                           // widths (<= 64): left operand, right operand
  const double *L; int ldl; int lunit;  // left operand (ml x n1), unit lower when lunit
  double *R; int ldr; int runit;        // right operand (ml x n2), unit lower when runit; the update's C
  const double *T1; int ldt1;           // n1 x n1 upper
  const double *T2; int ldt2;           // n2 x n2 upper (join)
  double *T12; int ldt12;               // join: the n1 x n2 block
  int rows;                             // rows per workgroup
  double *part, *w;                     // partial slots, the 64 x 64 middle factor
  unsigned *bar, *bar_next;             // this launch's barrier set, the next launch's
  int *err;
};

// one chunk of rows [c0, c0 + nr) of an operand, zero past nr rows and
// past ncols columns: loaded into registers (16 per thread: row t & 63 of
// columns t / 64 + 4 s, coalesced down the columns), then stored into LDS
// (RC x 64, column stride LDR); unit: the stored matrix is read as unit
// lower (global row c0 + r against column c)
__device__ __forceinline__ void chunk_load(double (&v)[16], const double *X, int ld, int c0, int nr, int ncols) {
  const int t = threadIdx.x, r = t & 63, cq = t >> 6;
#pragma unroll
  for (int s = 0; s < 16; s++) {
    const int c = cq + 4 * s;
    v[s] = (r < nr && c < ncols) ? X[(size_t)(c0 + r) + (size_t)c * ld] : 0.0;
  }
}
__device__ __forceinline__ void chunk_store(double *S, const double (&v)[16], int c0, int nr, int ncols, int unit) {
  const int t = threadIdx.x, r = t & 63, cq = t >> 6;
#pragma unroll
  for (int s = 0; s < 16; s++) {
    const int c = cq + 4 * s, gr = c0 + r;
    double x = v[s];
    if (unit && r < nr && c < ncols) x = (gr < c) ? 0.0 : (gr == c ? 1.0 : x);
    S[r + c * LDR] = x;
  }
}

// LDS 64 x 64 product on the four waves: O = op(A) B with A, B in LDS
// (column stride LDW, zero outside their triangles/extents); transA reads A
// transposed.  Returns each wave's 16 x 64 row block in MFMA output layout
// (tile tj, value v: row 16 wave + lane / 16 + 4 v, column 16 tj + lane % 16).
__device__ __forceinline__ void mm64(const double *A, bool transA, const double *B, d4_t (&o)[4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int tj = 0; tj < 4; tj++) o[tj] = d4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int s = 0; s < NW / 4; s++) {
    const int k = 4 * s + lk, i = 16 * wave + li;
    const double av = transA ? A[k + i * LDW] : A[i + k * LDW];
#pragma unroll
    for (int tj = 0; tj < 4; tj++)
      o[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, B[k + (16 * tj + li) * LDW], o[tj], 0, 0, 0);
  }
}
__device__ __forceinline__ void put64(double *O, const d4_t (&o)[4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int tj = 0; tj < 4; tj++)
#pragma unroll
    for (int v = 0; v < 4; v++) O[(16 * wave + lk + 4 * v) + (16 * tj + li) * LDW] = o[tj][v];
}
// an upper triangular n x n matrix (global, ld) into LDS, zero elsewhere
__device__ __forceinline__ void load_upper(double *S, const double *X, int ld, int n) {
  const int t = threadIdx.x, r = t & 63;
  double v[16];
#pragma unroll
  for (int s = 0; s < 16; s++) {
    const int c = (t >> 6) + 4 * s;
    v[s] = (r <= c && c < n) ? X[(size_t)r + (size_t)c * ld] : 0.0;
  }
#pragma unroll
  for (int s = 0; s < 16; s++) S[r + ((t >> 6) + 4 * s) * LDW] = v[s];
}

template <bool JOIN>
__global__ __launch_bounds__(NT) void node64_kernel(NodeArgs a) {
  extern __shared__ double smem[];
  double *Ls = smem + LDS_L, *Rs = smem + LDS_R, *Ws = smem + LDS_W, *Ts = smem + LDS_T;
  double *Us = smem + LDS_U, *red = smem + LDS_RED, *xs = smem + LDS_X;
  int *flag = (int *)(smem + LDS_FLAG);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = blockIdx.x, G = gridDim.x;
  const int li = lane & 15, lk = lane >> 4;
  if (g == 0 && tid == 0) {
    reset_barrier(a.bar_next);
    __hip_atomic_store(a.bar_next + BAR_STRIDE * (BAR_GROUPS + 1), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }

  // ---- 1. partial product P_g = L_g^T R_g (n1 x n2) over this workgroup's rows
  const int r_lo = g * a.rows, r_hi = min(a.ml, r_lo + a.rows);
  d4_t acc[4];
#pragma unroll
  for (int tj = 0; tj < 4; tj++) acc[tj] = d4_t{0.0, 0.0, 0.0, 0.0};
  const int nchunks = (r_hi - r_lo + RC - 1) / RC;
  double vl[16], vr[16];
  chunk_load(vl, a.L, a.ldl, r_lo, min(RC, r_hi - r_lo), a.n1);
  chunk_load(vr, a.R, a.ldr, r_lo, min(RC, r_hi - r_lo), a.n2);
  for (int c0 = r_lo; c0 < r_hi; c0 += RC) {
    const int nr = min(RC, r_hi - c0);
    if (c0 > r_lo) __syncthreads();  // the previous chunk's reads are done
    chunk_store(Ls, vl, c0, nr, a.n1, a.lunit);
    chunk_store(Rs, vr, c0, nr, a.n2, a.runit);
    __syncthreads();
    if (c0 + RC < r_hi) {  // the next chunk's loads in flight during this one's products
      chunk_load(vl, a.L, a.ldl, c0 + RC, min(RC, r_hi - c0 - RC), a.n1);
      chunk_load(vr, a.R, a.ldr, c0 + RC, min(RC, r_hi - c0 - RC), a.n2);
    }
#pragma unroll 4
    for (int s = 0; s < RC / 4; s++) {
      const int k = 4 * s + lk;
      const double av = Ls[k + (16 * wave + li) * LDR];
#pragma unroll
      for (int tj = 0; tj < 4; tj++)
        acc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Rs[k + (16 * tj + li) * LDR], acc[tj], 0, 0, 0);
    }
  }
  // T1 into LDS for the workgroups that sum columns
  if (g < a.n2) load_upper(Ts, a.T1, a.ldt1, a.n1);
  // the partial through LDS (Ws), then one coalesced 64-double run per column
  put64(Ws, acc);
  __syncthreads();
  for (int e = tid; e < NW * a.n2; e += NT) {
    const int j = e >> 6, i = e & 63;
    st_coh(&a.part[((size_t)j * G + g) * NW + i], Ws[i + j * LDW]);
  }
  grid_sync(a.bar, (unsigned)G, a.err);

  // ---- 2. columns j = g, g + G, ...: x = the sum of the G slots (each wave
  // a quarter of them in one batch of loads, then the four quarters in
  // order), times T1^T (update) or T1 (join)
  for (int j = g; j < a.n2; j += G) {
    const double *pj = a.part + (size_t)j * G * NW + lane;
    double v[GMAX / 4];
#pragma unroll
    for (int u = 0; u < GMAX / 4; u++) v[u] = (wave + 4 * u < G) ? ld_coh(pj + (size_t)(wave + 4 * u) * NW) : 0.0;
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < GMAX / 4; u++) s += v[u];
    red[wave * NW + lane] = s;
    __syncthreads();
    if (tid < NW) xs[tid] = ((red[tid] + red[NW + tid]) + red[2 * NW + tid]) + red[3 * NW + tid];
    __syncthreads();
    if (tid < NW) {
      double y = 0.0;
      if (JOIN) {
#pragma unroll 8
        for (int k = 0; k < NW; k++) y = fma(Ts[tid + k * LDW], xs[k], y);  // (T1 x)_i
      } else {
#pragma unroll 8
        for (int k = 0; k < NW; k++) y = fma(Ts[k + tid * LDW], xs[k], y);  // (T1^T x)_i
      }
      st_coh(&a.w[tid + j * NW], y);
    }
    __syncthreads();  // red, xs reused by the next column
  }

  if (!JOIN) {
    // ---- 3. update: after the second barrier, C_g -= V_g W
    grid_sync(a.bar, 2u * (unsigned)G, a.err);
#pragma unroll
    for (int s = 0; s < 16; s++) {
      const int e = tid + NT * s, c = e >> 6, r = e & 63;
      Ws[r + c * LDW] = (c < a.n2) ? ld_coh(&a.w[r + c * NW]) : 0.0;
    }
    d4_t o[4];
    for (int c0 = r_lo; c0 < r_hi; c0 += RC) {
      const int nr = min(RC, r_hi - c0);
      if (nchunks > 1) {  // the last chunk of step 1 is not the one needed: stage again
        chunk_load(vl, a.L, a.ldl, c0, nr, a.n1);
        chunk_load(vr, a.R, a.ldr, c0, nr, a.n2);
        __syncthreads();
        chunk_store(Ls, vl, c0, nr, a.n1, a.lunit);
        chunk_store(Rs, vr, c0, nr, a.n2, a.runit);
      }
      __syncthreads();
#pragma unroll
      for (int tj = 0; tj < 4; tj++)
#pragma unroll
        for (int v = 0; v < 4; v++) o[tj][v] = Rs[(16 * wave + lk + 4 * v) + (16 * tj + li) * LDR];
#pragma unroll 4
      for (int s = 0; s < NW / 4; s++) {
        const int k = 4 * s + lk;
        const double av = -Ls[(16 * wave + li) + k * LDR];
#pragma unroll
        for (int tj = 0; tj < 4; tj++)
          o[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Ws[k + (16 * tj + li) * LDW], o[tj], 0, 0, 0);
      }
      __syncthreads();  // every wave's reads of Rs are done
#pragma unroll
      for (int tj = 0; tj < 4; tj++)
#pragma unroll
        for (int v = 0; v < 4; v++) Rs[(16 * wave + lk + 4 * v) + (16 * tj + li) * LDR] = o[tj][v];
      __syncthreads();
      const int r = tid & 63;
#pragma unroll
      for (int s = 0; s < 16; s++) {
        const int c = (tid >> 6) + 4 * s;
        if (r < nr && c < a.n2) a.R[(size_t)(c0 + r) + (size_t)c * a.ldr] = Rs[r + c * LDR];
      }
    }
  } else {
    // ---- 3. join: the last workgroup out of step 2 forms T12 = -H T2
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores of H performed
    __syncthreads();
    if (tid == 0) {
      const unsigned old = __hip_atomic_fetch_add(a.bar + BAR_STRIDE * (BAR_GROUPS + 1), 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
      *flag = (old + 1 == (unsigned)G) ? 1 : 0;
    }
    __syncthreads();
    if (!*flag) return;
#pragma unroll
    for (int s = 0; s < 16; s++) {
      const int e = tid + NT * s, c = e >> 6, r = e & 63;
      Ws[r + c * LDW] = (c < a.n2) ? ld_coh(&a.w[r + c * NW]) : 0.0;  // H (n1 x n2), zero past n2
    }
    load_upper(Us, a.T2, a.ldt2, a.n2);
    __syncthreads();
    d4_t o[4];
    mm64(Ws, false, Us, o);  // H T2
    __syncthreads();
    put64(Ts, o);
    __syncthreads();
    const int r = tid & 63;
#pragma unroll
    for (int s = 0; s < 16; s++) {
      const int c = (tid >> 6) + 4 * s;
      if (r < a.n1 && c < a.n2) a.T12[(size_t)r + (size_t)c * a.ldt12] = -Ts[r + c * LDW];
    }
  }
}

struct NodeSetup {
  int state = 0;  // 0 not yet, 1 ready, -1 unavailable
  int ncu = 0, coop = 0;
  int rows = RC;  // target rows per workgroup (LILA_NODE_ROWS)
};
NodeSetup setup;

int node_ready(void) {
  if (setup.state) return setup.state;
  const void *ks[2] = {(const void *)node64_kernel<false>, (const void *)node64_kernel<true>};
  for (int i = 0; i < 2; i++) {
    if (hipFuncSetAttribute(ks[i], hipFuncAttributeMaxDynamicSharedMemorySize, LDS_BYTES) != hipSuccess) {
      (void)hipGetLastError();
      fprintf(stderr, "lila_hip: the node kernels cannot take %d bytes of LDS; nodes take the GEMM path\n",
              LDS_BYTES);
      setup.state = -1;
      return -1;
    }
  }
  int dev = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&setup.ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ks[0], NT, LDS_BYTES) != hipSuccess) {
    (void)hipGetLastError();
    setup.state = -1;
    return -1;
  }
  const char *ev = getenv("LILA_PANEL_COOP");
  setup.coop = (ev && atoi(ev) == 1) || per_cu < 1;
  ev = getenv("LILA_NODE_ROWS");
  if (ev && atoi(ev) >= 16 && atoi(ev) <= 4096) setup.rows = atoi(ev);
  setup.state = setup.ncu > 0 ? 1 : -1;
  return setup.state;
}

int node_launch(bool join, NodeArgs &a, const char *name, double flops, double bytes) {
  // G workgroups of a.rows rows: at most one per CU, chunks of 64 rows, and
  // at least min(64, ml / 16) workgroups so that the column sums of step 2
  // spread over many workgroups
  const int ncu = setup.ncu < GMAX ? setup.ncu : GMAX;
  int G = (a.ml + setup.rows - 1) / setup.rows;
  const int gmin = (a.ml + 15) / 16 < NW ? (a.ml + 15) / 16 : NW;
  if (G < gmin) G = gmin;
  if (G > ncu) G = ncu;
  if (G < 1) G = 1;
  a.rows = (a.ml + G - 1) / G;
  G = (a.ml + a.rows - 1) / a.rows;
  static int parity = 0;
  double *scr = ld_node_scratch();
  unsigned *bars = (unsigned *)(scr + OFF_BAR);
  a.part = scr + OFF_PART;
  a.w = scr + OFF_W;
  a.bar = bars + BAR_WORDS * parity;
  a.bar_next = bars + BAR_WORDS * (1 - parity);
  parity ^= 1;
  a.err = ld_device_flags();
  void *args[] = {&a};
  const void *k = join ? (const void *)node64_kernel<true> : (const void *)node64_kernel<false>;
  const int ph = ld_prof_begin(name);
  hipError_t e;
  if ((setup.coop || ld_lanes() > 1) && G > 1)
    e = hipLaunchCooperativeKernel(k, dim3(G), dim3(NT), args, (unsigned)LDS_BYTES, (hipStream_t)ld_stream());
  else
    e = hipLaunchKernel(k, dim3(G), dim3(NT), args, LDS_BYTES, (hipStream_t)ld_stream());
  ld_prof_end(ph, flops, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fprintf(stderr, "lila_hip: %s launch failed (G=%d rows=%d): %s\n", name, G, a.rows, hipGetErrorString(e));
    ld_set_info(-1);
  }
  return 0;
}

}  // namespace

extern "C" size_t ld_node_scratch_doubles(void) { return NODE_SCRATCH; }

extern "C" int ld_node_update64(int ml, int n1, int n2, const double *V, int ldv, const double *T1, int ldt,
                                double *C, int ldc) {
  if (ld_failed()) return 0;  // an earlier allocation of this call failed
  if (n1 < 1 || n2 < 1 || n1 > NW || n2 > NW || ml < n1 || node_ready() < 0) return -1;
  NodeArgs a;
  memset(&a, 0, sizeof a);
  a.ml = ml; a.n1 = n1; a.n2 = n2;
  a.L = V; a.ldl = ldv; a.lunit = 1;
  a.R = C; a.ldr = ldc; a.runit = 0;
  a.T1 = T1; a.ldt1 = ldt;
  return node_launch(false, a, "node_update64_kernel", 4.0 * ml * n1 * n2 + 2.0 * n1 * n1 * n2,
                     8.0 * ((double)ml * n1 + 2.0 * ml * n2));
}

extern "C" int ld_node_join64(int ml2, int n1, int n2, const double *V1, int ldv1, const double *V2, int ldv2,
                              const double *T1, int ldt1, const double *T2, int ldt2, double *T12, int ldt12) {
  if (ld_failed()) return 0;
  if (n1 < 1 || n2 < 1 || n1 > NW || n2 > NW || ml2 < n2 || node_ready() < 0) return -1;
  NodeArgs a;
  memset(&a, 0, sizeof a);
  a.ml = ml2; a.n1 = n1; a.n2 = n2;
  a.L = V1; a.ldl = ldv1; a.lunit = 0;
  a.R = (double *)V2; a.ldr = ldv2; a.runit = 1;
  a.T1 = T1; a.ldt1 = ldt1;
  a.T2 = T2; a.ldt2 = ldt2;
  a.T12 = T12; a.ldt12 = ldt12;
  return node_launch(true, a, "node_join64_kernel",
                     2.0 * ml2 * n1 * n2 + 2.0 * n1 * n1 * n2 + 2.0 * n1 * n2 * n2,
                     8.0 * (double)ml2 * (n1 + n2));
}
