// Inspector kernels: one pass over the vertices (VC) or edges (EC) applying
// the functor's filter, writing the frontier bitmaps with one 64-bit ballot
// per wavefront and reducing the workload counters.
//
// Behaviour follows __inspect_VC / __inspect_EC
// (src/abstraction/inspector.cuh:10-110). The loop is over 64-vertex words,
// one word per wavefront per iteration, so each word is written by exactly
// one lane and needs no atomics.
#pragma once
#include "../data_structures/active_set.h"
#include "../data_structures/functor.h"
#include "../utils/wave.h"
#include "mlp.h"

namespace gsw {

// Workgroup reduction, then one atomic into this workgroup's partial slot
// (slot blockIdx.x % STAT_SLOTS, a 128-byte line of its own); k_stat_finish
// reduces the slots.
__device__ __forceinline__ long long *stat_slot(long long *slots) {
  return slots + (int64_t)(blockIdx.x & (STAT_SLOTS - 1)) * STAT_STRIDE;
}
template <typename T>
__device__ __forceinline__ void flush_sum(T x, T *sh, long long *slot) {
  T r = block_sum(x, sh);
  if (threadIdx.x == 0 && r) atomicAdd((unsigned long long *)slot, (unsigned long long)r);
}
template <typename T>
__device__ __forceinline__ void flush_max(T x, T *sh, long long *slot) {
  T r = block_max(x, sh);
  if (threadIdx.x == 0 && r) atomicMax((long long *)slot, (long long)r);
}

template <typename G, typename F>
__global__ void __launch_bounds__(BLOCK)
    k_inspect_vc(active_set_t as, G g, F f, long long *slots) {
  __shared__ long long sh[BLOCK / WAVE];
  const int lane = lane_id();
  const int64_t n = g.nvertexs;
  const int64_t nw = (n + 63) >> 6;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  long long act = 0, inact = 0, pushw = 0, pullw = 0, maxa = 0, maxi = 0;
  for (int64_t w = wave; w < nw; w += nwaves) {
    const int64_t v = (w << 6) + lane;
    bool a = false, ni = false;
    if (v < n) {
      Status s = f.filter((int)v, g);
      a = (s == Active);
      ni = (s != Inactive);
      if (a) {
        long long od = g.get_out_degree((int)v);
        ++act;
        pushw += od;
        maxa = od > maxa ? od : maxa;
      }
      if (!ni) {
        long long id = g.get_in_degree((int)v);
        ++inact;
        pullw += id;
        maxi = id > maxi ? id : maxi;
      }
    }
    const uint64_t am = ballot(a), nm = ballot(ni);
    if (lane == 0) {
      as.bitmap.active.dg_bits[w] = am;
      as.bitmap.inactive.dg_bits[w] = nm;
    }
  }
  long long *slot = stat_slot(slots);
  flush_sum(act, sh, slot + 0);
  flush_sum(inact, sh, slot + 1);
  flush_sum(pushw, sh, slot + 2);
  flush_sum(pullw, sh, slot + 3);
  flush_max(maxa, sh, slot + 4);
  flush_max(maxi, sh, slot + 5);
}

// Edge-centric inspector: an edge stays active while filter(v0, v1) says so.
template <typename G, typename F>
__global__ void __launch_bounds__(BLOCK)
    k_inspect_ec(active_set_t as, G g, F f, long long *slots) {
  __shared__ long long sh[BLOCK / WAVE];
  const int lane = lane_id();
  const int64_t m = g.nedges;
  const int64_t nw = (m + 63) >> 6;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  long long act = 0;
  for (int64_t w = wave; w < nw; w += nwaves) {
    const int64_t e = (w << 6) + lane;
    bool next = false;
    if (e < m && as.bitmap.is_active((int)e)) {
      packed_t vv = g.dg_coo[e];
      int v0 = (int)(vv >> 32);
      int v1 = (int)(vv & 0xffffffffll);
      next = (f.filter(v0, v1, (typename G::edge_t *)nullptr) == Active);
    }
    const uint64_t am = ballot(next);
    if (lane == 0) as.bitmap.active.dg_bits[w] = am;
    act += next ? 1 : 0;
  }
  flush_sum(act, sh, stat_slot(slots));
}

}  // namespace gsw
