// gridsync.h -- cross-workgroup exchange and grid barriers for persistent
// launches whose workgroups are all resident (one per CU, grid <= CUs, or a
// cooperative launch), on gfx950.
//
// Values move between workgroups only through agent-scope atomics (relaxed
// FP64 adds and max, coherent loads and stores), which are performed at the
// coherence point, so no cache writeback or invalidate fences are needed.
#pragma once
#include <hip/hip_runtime.h>

namespace gridsync {

static __device__ __forceinline__ void st_coh(double *p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long *>(p),
                     (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ double ld_coh(const double *p) {
  unsigned long long u = __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)u);
}
static __device__ __forceinline__ void add_coh(double *p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ void max_coh(double *p, double v) {  // v >= 0
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
// The arrivals go through a two-level tree: atomics on one counter are
// performed one after another, a few hundred workgroups on one counter cost
// microseconds.  Workgroup g adds to the counter of its group g % 8, each
// counter 4 KiB from the others (different memory channels); the group's
// last arriver of the round (each round adds exactly the group's size, and
// a round's arrivals all precede the next round's) adds one to bar[0],
// which the waiters poll for one arrival per nonempty group and round.
// The last workgroup out of a launch resets them all (exit protocols).
// grid_sync in two halves, so a workgroup can work between its arrival and
// the wait (LDS-only work: its global writes before arrive are the ones
// the others see)
constexpr int BAR_GROUPS = 8;
constexpr int BAR_STRIDE = 1024;  // unsigned ints between two counters (4 KiB)
static __device__ void grid_arrive(unsigned *bar) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned G = gridDim.x, k = blockIdx.x % BAR_GROUPS;
    const unsigned gs = (G - k + BAR_GROUPS - 1) / BAR_GROUPS;  // workgroups of group k
    const unsigned old = __hip_atomic_fetch_add(&bar[BAR_STRIDE * (k + 1)], 1u, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
    if ((old + 1) % gs == 0) __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}
// target: the barrier's round times the grid size (rounds count from 1)
static __device__ void grid_wait(unsigned *bar, unsigned target, int *err) {
  if (threadIdx.x == 0) {
    const unsigned G = gridDim.x, groups = G < BAR_GROUPS ? G : BAR_GROUPS;
    const unsigned want = (target / G) * groups;
    unsigned spins = 0;
    while (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > (1u << 24)) {
        atomicOr(err, 1);
        break;
      }
    }
  }
  __syncthreads();
}

static __device__ void grid_sync(unsigned *bar, unsigned target, int *err) {
  grid_arrive(bar);
  grid_wait(bar, target, err);
}

// the exit protocols' reset of the barrier counters (last workgroup out)
static __device__ void reset_barrier(unsigned *bar) {
  for (int k = 0; k <= BAR_GROUPS; k++)
    __hip_atomic_store(&bar[BAR_STRIDE * k], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace gridsync
