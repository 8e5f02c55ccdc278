// device.h -- wave-level helpers shared by the persistent solve kernels
// (executors.hip: work units; region_exec.hip: regions).
#pragma once

#include "common.h"

namespace symamd {

// Optional per-item trace (100 MHz clock): trace[3u + slot] for item u
// (work unit or region task) when it fits the buffer; lane 0 writes.
__device__ __forceinline__ void trace_mark(uint64_t *trace, int64_t cap,
                                           int u, int slot) {
  if (trace && (threadIdx.x & 63) == 0 && 3 * (int64_t)u + 2 < cap)
    trace[3 * (int64_t)u + slot] = wall_ticks();
}

// The mark: a signalling-NaN bit pattern that no arithmetic produces
// (operations return quiet NaNs).  In the marked protocols a unit (or region task)
// overwrites its rows' right-hand sides with it right after reading them,
// so a solved value can be told apart from one not yet stored.
__device__ __forceinline__ bool is_mark(double v) {
  return __double_as_longlong(v) == (long long)0x7FF4D0D0CAFEF00Dull;
}
__device__ __forceinline__ bool is_mark(float v) {
  return __float_as_uint(v) == 0x7FA0CAFEu;
}
template <typename T>
__device__ __forceinline__ T mark_value();
template <>
__device__ __forceinline__ double mark_value<double>() {
  return __longlong_as_double((long long)0x7FF4D0D0CAFEF00Dull);
}
template <>
__device__ __forceinline__ float mark_value<float>() {
  return __uint_as_float(0x7FA0CAFEu);
}

// LDS writes of the wave are complete and visible to all its lanes.
__device__ __forceinline__ void wave_lds_sync() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_wave_barrier();
}

// Wait until every item u' of list[0 .. count) (global memory or LDS)
// carries `target` in flags[u' * mul + add].  One lane per item, one
// coherent load each; later rounds re-load only the pending ones after a
// short sleep.  A wait longer than `timeout` ticks of the 100 MHz clock
// records an error in state[ST_ERROR] and gives up, so every wave exits.
// (Batching several tags per lane into one round trip measured slower on the
// 128^3 IC(0) factor: the extra address setup sits on the critical path of
// every unit.)
// Items outside [0, limit) are read as item 0.
__device__ __forceinline__ void wait_list(const int *list, int count, const int *flags,
                                          int target, int *state, uint64_t timeout,
                                          int mul = 1, int add = 0, int limit = INT_MAX) {
  const int lane = threadIdx.x & 63;
  for (int base = 0; base < count; base += 64) {
    const int t = base + lane;
    bool pending = t < count;
    int item = pending ? list[t] : 0;
    item = (unsigned)item < (unsigned)limit ? item : 0;
    const int *f = pending ? &flags[item * mul + add] : flags;
    if (pending) pending = ld_dev(f) != target;
    if (!__any(pending)) continue;
    const uint64_t t0 = wall_ticks();
    int rounds = 0;
    while (__any(pending)) {
      __builtin_amdgcn_s_sleep(1);
      if (pending) pending = ld_dev(f) != target;
      if ((++rounds & 63) == 0 && wall_ticks() - t0 > timeout) {
        if (pending) atomicOr(&state[ST_ERROR], 1);
        pending = false;
      }
    }
  }
  asm volatile("" ::: "memory");
}

// wait_list over the dependencies of item u, dlist[dptr[u] .. dptr[u+1]).
__device__ __forceinline__ void wait_units(const int *__restrict__ dptr,
                                           const int *__restrict__ dlist, int u,
                                           const int *flags, int target, int *state,
                                           uint64_t timeout) {
  const int d0 = dptr[u];
  wait_list(dlist + d0, dptr[u + 1] - d0, flags, target, state, timeout);
}

// Next item of this wave: round-robin over a co-resident grid (cursor
// advances by the number of waves) or the next ticket.
__device__ __forceinline__ int next_unit(int *state, int ticketed, int &cursor,
                                         int stride) {
  if (ticketed) {
    int u = 0;
    if ((threadIdx.x & 63) == 0) u = atomicAdd(&state[ST_TICKET], 1);
    return __shfl(u, 0);
  }
  const int u = cursor;
  cursor += stride;
  return u;
}

}  // namespace symamd
