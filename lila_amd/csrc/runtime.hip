// runtime.hip -- library state for the LILA MI355X path: stream, device
// workspace arena, status slot, pointer classification, and the small
// element-wise kernels the drivers need (copies, fills, sign flips,
// generators).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "lila_dev.h"

extern "C" size_t ld_panel_scratch_doubles(void);
extern "C" size_t ld_node_scratch_doubles(void);

#define HIPCHECK(x)                                                           \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      fprintf(stderr, "lila_hip: %s failed: %s (%s:%d)\n", #x,                  \
              hipGetErrorString(e_), __FILE__, __LINE__);                       \
      ld_set_info(-1);                                                          \
    }                                                                           \
  } while (0)

namespace {
// Workspace arena: a list of device chunks used as one bump stack.  A chunk
// is never freed while a call tree runs (pointers handed out stay valid);
// ld_ws_consolidate() folds several chunks into one after a synchronize.
constexpr int MAXCHUNK = 32;
struct Chunk {
  char *p;
  size_t cap, top;
};
struct Arena {
  Chunk ch[MAXCHUNK];
  int nch = 0, cur = 0;
};
// Two lanes of work: lane 0 is the library stream (lila_set_stream), lane 1
// an internal stream for work that may overlap lane 0's (the drivers fork
// and join it with events).  Each lane bump-allocates from its own arena, so
// the stream order of a lane is what keeps its workspace reuse safe.
constexpr int NEV = 32;
struct State {
  int inited = 0;
  hipStream_t stream = nullptr;       // the current lane's stream
  hipStream_t user_stream = nullptr;  // lane 0
  hipStream_t aux_stream = nullptr;   // lane 1
  int lane = 0, lanes = 1;
  Arena ar[2];
  hipEvent_t ev[NEV];
  int evi = 0;
  int info = 0;
  int *dflags = nullptr;  // device flags: [0] barrier timeout, [1] potrf info, [8..9] barrier
  double *pscratch = nullptr;  // panel reduction vectors (kept zero between launches)
  double *nscratch = nullptr;  // node kernels' partial slots and barrier counters
  ld_tune_t tune = {384, 128, 512, 64, 64, 128, 256, 1, 0, 128, 128, 0, 64, 0, 64, 0, 0, 1, 0, 1, 0};
  int failed = 0;  // a workspace allocation of the current call failed
};
State g;
}  // namespace

extern "C" {

int ld_init(void) {
  if (g.inited) return 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    fprintf(stderr, "lila_hip: no HIP device visible\n");
    return -1;
  }
  if (hipMalloc(&g.dflags, 64 * sizeof(int)) != hipSuccess) return -1;
  if (hipMemset(g.dflags, 0, 64 * sizeof(int)) != hipSuccess) return -1;
  {
    const size_t pb = ld_panel_scratch_doubles() * sizeof(double);
    if (hipMalloc(&g.pscratch, pb) != hipSuccess) return -1;
    if (hipMemset(g.pscratch, 0, pb) != hipSuccess) return -1;
    const size_t nb = ld_node_scratch_doubles() * sizeof(double);
    if (hipMalloc(&g.nscratch, nb) != hipSuccess) return -1;
    if (hipMemset(g.nscratch, 0, nb) != hipSuccess) return -1;
  }
  if (hipStreamCreateWithFlags(&g.aux_stream, hipStreamNonBlocking) != hipSuccess) return -1;
  for (int k = 0; k < NEV; k++)
    if (hipEventCreateWithFlags(&g.ev[k], hipEventDisableTiming) != hipSuccess) return -1;
  const char *e = getenv("LILA_LANES");
  if (e) g.lanes = atoi(e) == 2 ? 2 : 1;
  e = getenv("LILA_PANEL_ROWS");
  if (e) g.tune.panel_rows_per_wg = g.tune.panel_rows_cqr = atoi(e);
  e = getenv("LILA_GEMM_MIN_WGS");
  if (e) g.tune.gemm_min_wgs = atoi(e);
  e = getenv("LILA_GEMM_TILE");  // "BMxBN", e.g. 128x64
  if (e) {
    int bm = 0, bn = 0;
    if (sscanf(e, "%dx%d", &bm, &bn) == 2 && (bm == 32 || bm == 64 || bm == 128) &&
        (bn == 32 || bn == 64 || bn == 128)) {
      g.tune.gemm_bm = bm;
      g.tune.gemm_bn = bn;
    }
  }
  e = getenv("LILA_GEMM_TILE_T");  // the same for products with op(A) = A^T
  if (e) {
    int bm = 0, bn = 0;
    if (sscanf(e, "%dx%d", &bm, &bn) == 2 && (bm == 32 || bm == 64 || bm == 128) &&
        (bn == 32 || bn == 64 || bn == 128)) {
      g.tune.gemm_bm_t = bm;
      g.tune.gemm_bn_t = bn;
    }
  }
  e = getenv("LILA_GEMM_SMALL_TILES");
  if (e) g.tune.gemm_small_tiles = atoi(e);
  e = getenv("LILA_GEMM_SMALL_TILES_T");
  if (e) g.tune.gemm_small_tiles_t = atoi(e);
  e = getenv("LILA_GEMM_V2");
  if (e) g.tune.gemm_v2 = atoi(e);
  e = getenv("LILA_Q_AFTER");
  if (e) g.tune.q_after = atoi(e) ? 1 : 0;
  e = getenv("LILA_ROCBLAS");
  if (e) g.tune.gemm_rocblas = atoi(e) ? 1 : 0;
  e = getenv("LILA_NODE_FUSED");
  if (e) g.tune.node_fused = atoi(e) ? 1 : 0;
  e = getenv("LILA_GEMM_KMIN");
  if (e) g.tune.gemm_kmin = atoi(e);
  e = getenv("LILA_PANEL_MODE");
  if (e) g.tune.panel_mode = atoi(e) ? 1 : 0;
  e = getenv("LILA_PANEL_WIDTH");  // as lila_set_panel_width
  if (e && ((atoi(e) >= 8 && atoi(e) <= LD_PANEL_MAX) || atoi(e) == 64)) g.tune.panel_nb = atoi(e);
  // the static scratch of lila_dgels_batched_noalloc on the current device
  if (ld_batch_noalloc_setup() != 0) return -1;
  g.inited = 1;
  return 0;
}

ld_tune_t *ld_tune(void) { return &g.tune; }

void *ld_stream(void) { return (void *)g.stream; }
// The switch is ordered: the new stream waits for everything enqueued on
// the old one (library and device-pointer calls are asynchronous, and the
// workspace arena, the panel scratch and the barrier counters they share
// are kept safe by stream order only).
void ld_set_stream(void *s) {
  const hipStream_t ns = (hipStream_t)s;
  if (g.inited && ns != g.user_stream) {
    hipEvent_t e = g.ev[g.evi];
    g.evi = (g.evi + 1) % NEV;
    HIPCHECK(hipEventRecord(e, g.user_stream));
    HIPCHECK(hipEventRecord(g.ev[g.evi], g.aux_stream));
    HIPCHECK(hipStreamWaitEvent(ns, e, 0));
    HIPCHECK(hipStreamWaitEvent(ns, g.ev[g.evi], 0));
    g.evi = (g.evi + 1) % NEV;
  }
  g.user_stream = ns;
  if (g.lane == 0) g.stream = g.user_stream;
}

int ld_lanes(void) { return g.lanes; }
int ld_set_lanes(int n) {
  const int old = g.lanes;
  g.lanes = (n == 1) ? 1 : 2;
  return old;
}
void ld_set_lane(int l) {
  g.lane = l ? 1 : 0;
  g.stream = g.lane ? g.aux_stream : g.user_stream;
}
// lane 1 waits for everything lane 0 has enqueued so far
void ld_lane_fork(void) {
  hipEvent_t e = g.ev[g.evi];
  g.evi = (g.evi + 1) % NEV;
  HIPCHECK(hipEventRecord(e, g.user_stream));
  HIPCHECK(hipStreamWaitEvent(g.aux_stream, e, 0));
}
static hipError_t sync_lanes(void) {
  hipError_t e = hipStreamSynchronize(g.aux_stream);
  hipError_t f = hipStreamSynchronize(g.user_stream);
  return e != hipSuccess ? e : f;
}
// lane 0 waits for everything lane 1 has enqueued so far
void ld_lane_join(void) {
  hipEvent_t e = g.ev[g.evi];
  g.evi = (g.evi + 1) % NEV;
  HIPCHECK(hipEventRecord(e, g.aux_stream));
  HIPCHECK(hipStreamWaitEvent(g.user_stream, e, 0));
}

// the raw classification, for error messages: hipPointerGetAttributes'
// status and the memory type it reports
void ld_ptr_diag(const void *p, int *err, int *type) {
  hipPointerAttribute_t a;
  memset(&a, 0, sizeof a);
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) (void)hipGetLastError();
  *err = (int)e;
  *type = (int)a.type;
}

int ld_is_device_ptr(const void *p) {
  if (!p) return 0;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

static int ws_new_chunk(Arena &a, size_t bytes) {
  if (a.nch >= MAXCHUNK) {
    fprintf(stderr, "lila_hip: workspace chunk list full\n");
    ld_set_info(-3);
    return -1;
  }
  char *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    fprintf(stderr, "lila_hip: cannot allocate %zu bytes of workspace\n", bytes);
    ld_set_info(-2);
    return -1;
  }
  a.ch[a.nch].p = p;
  a.ch[a.nch].cap = bytes;
  a.ch[a.nch].top = 0;
  a.nch++;
  return 0;
}

// make sure at least `bytes` are free in one chunk above the current top
int ld_ws_reserve(size_t bytes) {
  Arena &a = g.ar[g.lane];
  bytes = (bytes + 255) & ~(size_t)255;
  if (a.nch > 0 && a.ch[a.cur].top + bytes <= a.ch[a.cur].cap) return 0;
  for (int c = a.cur + 1; c < a.nch; c++)
    if (a.ch[c].cap >= bytes) return 0;
  size_t want = bytes < ((size_t)64 << 20) ? ((size_t)64 << 20) : bytes;
  return ws_new_chunk(a, want);
}

// A workspace chunk that cannot be allocated returns NULL and marks the call
// failed (info -2): from then on every ld_* launcher and copy of the call
// returns at once, so no kernel sees the NULL, and the API call returns -2.
void *ld_ws_alloc(size_t bytes) {
  Arena &a = g.ar[g.lane];
  bytes = (bytes + 255) & ~(size_t)255;
  if (g.failed) return nullptr;
  if (a.nch == 0 && ws_new_chunk(a, bytes < ((size_t)64 << 20) ? ((size_t)64 << 20) : bytes) != 0) {
    g.failed = 1;
    return nullptr;
  }
  if (a.ch[a.cur].top + bytes > a.ch[a.cur].cap) {
    // move to the next chunk that fits (chunks above cur are empty)
    int c = a.cur + 1;
    while (c < a.nch && a.ch[c].cap < bytes) c++;
    if (c >= a.nch) {
      size_t want = 2 * a.ch[a.nch - 1].cap;
      if (want < bytes) want = bytes;
      if (ws_new_chunk(a, want) != 0) {
        g.failed = 1;
        return nullptr;
      }
      c = a.nch - 1;
    }
    a.cur = c;
    a.ch[c].top = 0;
  }
  void *p = a.ch[a.cur].p + a.ch[a.cur].top;
  a.ch[a.cur].top += bytes;
  return p;
}

// mark = (chunk << 40) | top
size_t ld_ws_mark(void) {
  const Arena &a = g.ar[g.lane];
  if (a.nch == 0) return 0;
  return ((size_t)a.cur << 40) | a.ch[a.cur].top;
}
void ld_ws_release(size_t mark) {
  Arena &a = g.ar[g.lane];
  if (a.nch == 0) return;
  int c = (int)(mark >> 40);
  size_t top = mark & (((size_t)1 << 40) - 1);
  for (int k = c + 1; k < a.nch; k++) a.ch[k].top = 0;
  a.cur = c;
  a.ch[c].top = top;
}

// after a synchronize, with nothing in use: fold each lane's chunks into one
void ld_ws_consolidate(void) {
  if (g.ar[0].nch <= 1 && g.ar[1].nch <= 1) return;
  HIPCHECK(hipStreamSynchronize(g.user_stream));
  HIPCHECK(hipStreamSynchronize(g.aux_stream));
  for (int l = 0; l < 2; l++) {
    Arena &a = g.ar[l];
    if (a.nch <= 1) continue;
    size_t tot = 0;
    for (int k = 0; k < a.nch; k++) tot += a.ch[k].cap;
    for (int k = 0; k < a.nch; k++) HIPCHECK(hipFree(a.ch[k].p));
    a.nch = 0;
    a.cur = 0;
    ws_new_chunk(a, tot);
  }
}

void *ld_dev_malloc(size_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) {
    (void)hipGetLastError();  // reported through the info slot, not left pending
    ld_set_info(-2);
    return nullptr;
  }
  return p;
}
void ld_dev_free(void *p) {
  if (p) HIPCHECK(hipFree(p));
}
void ld_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (bytes) HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g.stream));
}
void ld_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (bytes) {
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g.stream));
    HIPCHECK(hipStreamSynchronize(g.stream));
  }
}

// 1 when the library stream is being captured into a hipGraph
int ld_stream_capturing(void) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(g.stream, &st) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return st != hipStreamCaptureStatusNone;
}

int ld_failed(void) { return g.failed; }
void ld_clear_failed(void) { g.failed = 0; }

void ld_set_info(int info) {
  if (g.info == 0 || info < 0) g.info = info;
}
int ld_get_info(void) { return g.info; }
void ld_clear_info(void) { g.info = 0; }

int *ld_device_flags(void) { return g.dflags; }
double *ld_panel_scratch(void) { return g.pscratch; }
double *ld_node_scratch(void) { return g.nscratch; }

int ld_check_device_flags(void) {
  int h[4] = {0, 0, 0, 0};
  if (!g.dflags) return 0;
  HIPCHECK(hipMemcpyAsync(h, g.dflags, sizeof h, hipMemcpyDeviceToHost, g.stream));
  HIPCHECK(hipStreamSynchronize(g.stream));
  if (h[0]) {
    fprintf(stderr, "lila_hip: grid barrier timed out in a panel kernel\n");
    ld_set_info(-4);
  }
  if (h[1] > 0) ld_set_info(h[1]);
  if (h[0] || h[1]) HIPCHECK(hipMemsetAsync(g.dflags, 0, sizeof h, g.stream));
  return h[0] ? -4 : h[1];
}

int ld_sync(void) {
  hipError_t e = sync_lanes();
  if (e != hipSuccess) {
    fprintf(stderr, "lila_hip: stream error: %s\n", hipGetErrorString(e));
    ld_set_info(-1);
    return -1;
  }
  return ld_check_device_flags();
}

void ld_fatal(const char *msg) {
  fprintf(stderr, "lila_hip: %s\n", msg);
  ld_set_info(-5);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// opt-in launch profiling: a start/stop hipEvent pair on the library stream
// around each launch of a kernel class, with the launch's algorithmic flops
// and bytes; read back (after a synchronize) as per-class totals.
// ---------------------------------------------------------------------------
namespace {
struct ProfRec {
  int cls;
  hipEvent_t a, b;
  double flops, bytes;
};
struct ProfCls {
  char name[96];
  long launches;
  double flops, bytes, ms;
};
int prof_on = 0;
char prof_only[96] = "";  // when set, only classes with this exact name are recorded
ProfRec *prec = nullptr;
int nrec = 0, caprec = 0;
ProfCls pcls[64];
int ncls = 0;
hipEvent_t *evpool = nullptr;
int nev = 0, capev = 0, evused = 0;

hipEvent_t ev_get() {
  if (evused == nev) {
    if (nev == capev) {
      capev = capev ? 2 * capev : 4096;
      evpool = (hipEvent_t *)realloc(evpool, sizeof(hipEvent_t) * capev);
    }
    hipEventCreate(&evpool[nev++]);
  }
  return evpool[evused++];
}

void prof_flush() {
  for (int r = 0; r < nrec; r++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, prec[r].a, prec[r].b) == hipSuccess) {
      ProfCls &c = pcls[prec[r].cls];
      c.launches++;
      c.flops += prec[r].flops;
      c.bytes += prec[r].bytes;
      c.ms += ms;
    }
  }
  nrec = 0;
  evused = 0;
}
}  // namespace

extern "C" {

int ld_prof_begin(const char *name) {
  if (!prof_on) return -1;
  if (prof_only[0] && strcmp(prof_only, name) != 0) return -1;
  int c = 0;
  for (; c < ncls; c++)
    if (strcmp(pcls[c].name, name) == 0) break;
  if (c == ncls) {
    if (ncls == 64) return -1;
    memset(&pcls[c], 0, sizeof pcls[c]);
    snprintf(pcls[c].name, sizeof pcls[c].name, "%s", name);
    ncls++;
  }
  if (evused + 2 > 60000) {  // bound the outstanding events: drain
    sync_lanes();
    prof_flush();
  }
  if (nrec == caprec) {
    caprec = caprec ? 2 * caprec : 4096;
    prec = (ProfRec *)realloc(prec, sizeof(ProfRec) * caprec);
  }
  ProfRec &r = prec[nrec];
  r.cls = c;
  r.a = ev_get();
  r.b = ev_get();
  r.flops = r.bytes = 0;
  hipEventRecord(r.a, g.stream);
  return nrec++;
}

void ld_prof_end(int h, double flops, double bytes) {
  if (h < 0) return;
  prec[h].flops = flops;
  prec[h].bytes = bytes;
  hipEventRecord(prec[h].b, g.stream);
}

// ---- exported (include/lila.h extensions) ----
int lila_set_panel_mode(int mode) {
  if (ld_init() != 0) return -1;
  const int old = g.tune.panel_mode;
  g.tune.panel_mode = mode ? 1 : 0;
  return old;
}

int lila_set_panel_width(int nb) {
  if (ld_init() != 0) return -1;
  if ((nb < 8 || nb > LD_PANEL_MAX) && nb != 64) return -1;
  const int old = ((g.tune.panel_nb >= 8 && g.tune.panel_nb <= LD_PANEL_MAX) || g.tune.panel_nb == 64)
                      ? g.tune.panel_nb
                      : LD_PANEL_MAX;
  g.tune.panel_nb = nb;
  return old;
}

int lila_set_q_after(int on) {
  if (ld_init() != 0) return -1;
  const int old = g.tune.q_after;
  g.tune.q_after = on ? 1 : 0;
  return old;
}

int lila_set_gemm_backend(int b) {
  if (ld_init() != 0) return -1;
  if (b != 0 && b != 1) return -1;
  const int old = g.tune.gemm_rocblas;
  g.tune.gemm_rocblas = b;
  return old;
}

int lila_set_node_fused(int on) {
  if (ld_init() != 0) return -1;
  const int old = g.tune.node_fused;
  g.tune.node_fused = on ? 1 : 0;
  return old;
}

int lila_set_gram_precision(int prec) {
  if (ld_init() != 0) return -1;
  if (prec < 0 || prec > 3) return -1;
  const int old = g.tune.gram_prec;
  g.tune.gram_prec = prec;
  return old;
}

int lila_set_batched_grid(int wgs) {
  if (ld_init() != 0) return -1;
  if (wgs < 0) return -1;
  const int old = g.tune.batch_grid;
  g.tune.batch_grid = wgs;
  return old;
}

int lila_set_lanes(int n) {
  if (ld_init() != 0) return -1;
  sync_lanes();
  return ld_set_lanes(n);
}

int lila_set_gemm_lowp(int on) {
  if (ld_init() != 0) return -1;
  const int old = g.tune.gemm_lowp;
  g.tune.gemm_lowp = on ? 1 : 0;
  return old;
}

int lila_prof_enable(int on) {
  if (ld_init() != 0) return -1;
  sync_lanes();
  prof_flush();
  if (on) ncls = 0;
  prof_on = on;
  return 0;
}

// restrict recording to one kernel class (NULL or "" = all classes)
int lila_prof_only(const char *kernel_class) {
  if (ld_init() != 0) return -1;
  sync_lanes();
  prof_flush();
  snprintf(prof_only, sizeof prof_only, "%s", kernel_class ? kernel_class : "");
  return 0;
}

// "name|launches|flops|bytes|ms\n" per kernel class; returns the class count
int lila_prof_read(char *buf, int buflen) {
  if (ld_init() != 0) return -1;
  sync_lanes();
  prof_flush();
  int off = 0;
  for (int c = 0; c < ncls && off < buflen; c++)
    off += snprintf(buf + off, (size_t)(buflen - off), "%s|%ld|%.6e|%.6e|%.6f\n", pcls[c].name,
                    pcls[c].launches, pcls[c].flops, pcls[c].bytes, pcls[c].ms);
  return ncls;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// element-wise kernels
// ---------------------------------------------------------------------------
namespace {

__global__ void lacpy_kernel(int uplo, int m, int n, const double *__restrict__ A, int lda,
                             double *__restrict__ B, int ldb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= m || j >= n) return;
  if (uplo == 'U' && i > j) return;
  if (uplo == 'L' && i < j) return;
  B[(size_t)i + (size_t)j * ldb] = A[(size_t)i + (size_t)j * lda];
}

__global__ void laset_kernel(int uplo, int m, int n, double off, double dg, double *A, int lda) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= m || j >= n) return;
  if (i == j) { A[(size_t)i + (size_t)j * lda] = dg; return; }
  if (uplo == 'U' && i > j) return;
  if (uplo == 'L' && i < j) return;
  A[(size_t)i + (size_t)j * lda] = off;
}

__global__ void add_diag_kernel(int n, double *A, int lda, double v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(size_t)i + (size_t)i * lda] += v;
}

__global__ void set_diag_kernel(int n, const double *__restrict__ v, int incv, int recip, double *A,
                                int lda) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double x = v[(size_t)i * incv];
    A[(size_t)i + (size_t)i * lda] = recip ? 1.0 / x : x;
  }
}

__global__ void sub_kernel(int m, int n, const double *__restrict__ W, int ldw, double *C, int ldc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i < m && j < n) C[(size_t)i + (size_t)j * ldc] -= W[(size_t)i + (size_t)j * ldw];
}

__global__ void flip_cols_kernel(int m, int n, double *X, int ldx, const double *S) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i < m && j < n && S[j] == -1.0) X[(size_t)i + (size_t)j * ldx] = -X[(size_t)i + (size_t)j * ldx];
}

__global__ void flip_rows_upper_kernel(int n, double *X, int ldx, const double *S) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;  // row
  int c = blockIdx.y;                              // column
  if (k < n && c < n && c >= k && S[k] == -1.0)
    X[(size_t)k + (size_t)c * ldx] = -X[(size_t)k + (size_t)c * ldx];
}

__global__ void flip_rows_kernel(int nrows, int n, double *X, int ldx, const double *S) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int c = blockIdx.y;
  if (k < nrows && c < n && S[k] == -1.0) X[(size_t)k + (size_t)c * ldx] = -X[(size_t)k + (size_t)c * ldx];
}

__global__ void move_lower_kernel(int n, double *Tq, int ldt, double *A, int lda) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  int c = blockIdx.y;
  if (r < n && c < n && r > c) {
    A[(size_t)r + (size_t)c * lda] = Tq[(size_t)r + (size_t)c * ldt];
    Tq[(size_t)r + (size_t)c * ldt] = 0.0;
  }
}

__global__ void t2tau_kernel(int n, int i, int mt, const double *T, int ldt, double *tau) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) {
    int g = i + c;
    tau[c] = T[(size_t)(g % mt) + (size_t)g * ldt];
  }
}

// one workgroup per reflector: tau = 2 / (1 + ||A(g+1:m, g)||^2)
__global__ void v2tau_kernel(int m, int i, const double *A, int lda, double *tau) {
  __shared__ double red[256];
  int c = blockIdx.x, g = i + c;
  double s = 0.0;
  for (int r = g + 1 + threadIdx.x; r < m; r += blockDim.x) {
    double v = A[(size_t)r + (size_t)g * lda];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) tau[c] = 2.0 / (1.0 + red[0]);
}

__device__ __forceinline__ uint64_t sm64_at(uint64_t seed, uint64_t idx) {
  // value number idx (0-based) of the splitmix64 stream started at seed
  uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ void fill_uniform_kernel(uint64_t seed, int m, int n, double *A, int ld, double lo, double hi) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t tot = (size_t)m * n;
  if (e >= tot) return;
  int r = (int)(e % m), c = (int)(e / m);
  double u = (double)(sm64_at(seed, e) >> 11) * 0x1.0p-53;
  A[(size_t)r + (size_t)c * ld] = lo + (hi - lo) * u;
}

__global__ void fill_normal_kernel(uint64_t seed, int m, int n, double *A, int ld) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t tot = (size_t)m * n;
  if (e >= tot) return;
  int r = (int)(e % m), c = (int)(e / m);
  size_t pair = e & ~(size_t)1;
  double u1 = (double)((sm64_at(seed, pair) >> 11) + 1) * 0x1.0p-53;
  double u2 = (double)(sm64_at(seed, pair + 1) >> 11) * 0x1.0p-53;
  double rad = sqrt(-2.0 * log(u1));
  double th = 6.283185307179586 * u2;
  A[(size_t)r + (size_t)c * ld] = (e & 1) ? rad * sin(th) : rad * cos(th);
}

inline dim3 grid2(int m, int n) { return dim3((unsigned)((m + 255) / 256), (unsigned)n); }

}  // namespace

extern "C" {

void ld_lacpy(int uplo, int m, int n, const double *A, int lda, double *B, int ldb) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL(lacpy_kernel, grid2(m, n), dim3(256), 0, g.stream, uplo, m, n, A, lda, B, ldb);
}
void ld_laset(int uplo, int m, int n, double off, double dg, double *A, int lda) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL(laset_kernel, grid2(m, n), dim3(256), 0, g.stream, uplo, m, n, off, dg, A, lda);
}
void ld_add_diag(int n, double *A, int lda, double v) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (n <= 0) return;
  hipLaunchKernelGGL(add_diag_kernel, dim3((n + 255) / 256), dim3(256), 0, g.stream, n, A, lda, v);
}
void ld_set_diag(int n, const double *v, int incv, int recip, double *A, int lda) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (n <= 0) return;
  hipLaunchKernelGGL(set_diag_kernel, dim3((n + 255) / 256), dim3(256), 0, g.stream, n, v, incv, recip,
                     A, lda);
}
void ld_sub(int m, int n, const double *W, int ldw, double *C, int ldc) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL(sub_kernel, grid2(m, n), dim3(256), 0, g.stream, m, n, W, ldw, C, ldc);
}
void ld_flip_cols(int m, int n, double *X, int ldx, const double *S) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL(flip_cols_kernel, grid2(m, n), dim3(256), 0, g.stream, m, n, X, ldx, S);
}
void ld_flip_rows_upper(int n, double *X, int ldx, const double *S) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (n <= 0) return;
  hipLaunchKernelGGL(flip_rows_upper_kernel, grid2(n, n), dim3(256), 0, g.stream, n, X, ldx, S);
}
void ld_flip_rows(int nrows, int n, double *X, int ldx, const double *S) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (nrows <= 0 || n <= 0) return;
  hipLaunchKernelGGL(flip_rows_kernel, grid2(nrows, n), dim3(256), 0, g.stream, nrows, n, X, ldx, S);
}
void ld_move_lower(int n, double *Tq, int ldt, double *A, int lda) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (n <= 1) return;
  hipLaunchKernelGGL(move_lower_kernel, grid2(n, n), dim3(256), 0, g.stream, n, Tq, ldt, A, lda);
}
void ld_T2tau(int n, int i, int mt, const double *T, int ldt, double *tau) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (n <= 0) return;
  hipLaunchKernelGGL(t2tau_kernel, dim3((n + 255) / 256), dim3(256), 0, g.stream, n, i, mt, T, ldt, tau);
}
void ld_V2tau(int m, int n, int i, const double *A, int lda, double *tau) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  if (n <= 0) return;
  hipLaunchKernelGGL(v2tau_kernel, dim3(n), dim3(256), 0, g.stream, m, i, A, lda, tau);
}
void ld_fill_uniform(unsigned long long seed, int m, int n, double *A, int ld, double lo, double hi) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  size_t tot = (size_t)m * n;
  if (!tot) return;
  hipLaunchKernelGGL(fill_uniform_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, g.stream,
                     (uint64_t)seed, m, n, A, ld, lo, hi);
}
void ld_fill_normal(unsigned long long seed, int m, int n, double *A, int ld) {
  if (ld_failed()) return;  // an earlier allocation of this call failed
  size_t tot = (size_t)m * n;
  if (!tot) return;
  hipLaunchKernelGGL(fill_normal_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, g.stream,
                     (uint64_t)seed, m, n, A, ld);
}

}  // extern "C"
