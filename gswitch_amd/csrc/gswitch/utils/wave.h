// wave64 primitives for CDNA (gfx950).
//
// The reference's primitives are 32-lane (src/utils/intrinsics.cuh:6-233:
// _ballot, warpScan, blockReduceSum, __upper_bound ...). Here everything is
// built for 64-lane wavefronts: 64-bit ballots, DPP row-shift/row-broadcast
// scans, mbcnt for in-wave ranks, and LDS only across waves.
#pragma once
#include "common.h"

namespace gsw {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// popcount of `mask` restricted to lanes below the calling lane
__device__ __forceinline__ int rank_below(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi(
      (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ int readlane(int x, int lane) {
  return __builtin_amdgcn_readlane(x, lane);
}
__device__ __forceinline__ int first_lane_value(int x) {
  return __builtin_amdgcn_readfirstlane(x);
}

// Inclusive prefix sum across the 64 lanes: four DPP row shifts inside each
// 16-lane row, then row_bcast:15 and row_bcast:31 carry the row totals.
// Lanes that are inactive contribute nothing only if they hold 0; callers
// run this with the whole wave active.
__device__ __forceinline__ int wave_inclusive_scan(int x) {
  int s = x;
  s += __builtin_amdgcn_update_dpp(0, s, 0x111, 0xf, 0xf, false);  // row_shr:1
  s += __builtin_amdgcn_update_dpp(0, s, 0x112, 0xf, 0xf, false);  // row_shr:2
  s += __builtin_amdgcn_update_dpp(0, s, 0x114, 0xf, 0xf, false);  // row_shr:4
  s += __builtin_amdgcn_update_dpp(0, s, 0x118, 0xf, 0xf, false);  // row_shr:8
  s += __builtin_amdgcn_update_dpp(0, s, 0x142, 0xa, 0xf, false);  // row_bcast:15
  s += __builtin_amdgcn_update_dpp(0, s, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return s;
}

__device__ __forceinline__ int wave_exclusive_scan(int x, int &total) {
  int inc = wave_inclusive_scan(x);
  total = readlane(inc, 63);
  return inc - x;
}

__device__ __forceinline__ int wave_sum(int x) {
  return readlane(wave_inclusive_scan(x), 63);
}

template <typename T> __device__ __forceinline__ T wave_sum_any(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

template <typename T> __device__ __forceinline__ T wave_max_any(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T y = __shfl_xor(x, o, 64);
    x = x > y ? x : y;
  }
  return x;
}

// Block-wide sum / max of a per-thread value; result valid in thread 0.
// `sh` must hold blockDim.x / 64 elements.
template <typename T>
__device__ __forceinline__ T block_sum(T x, T *sh) {
  x = wave_sum_any(x);
  const int w = threadIdx.x >> LOG_WAVE;
  const int nw = (blockDim.x + WAVE - 1) >> LOG_WAVE;
  __syncthreads();
  if (lane_id() == 0) sh[w] = x;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < nw; ++i) r += sh[i];
  return r;
}

template <typename T>
__device__ __forceinline__ T block_max(T x, T *sh) {
  x = wave_max_any(x);
  const int w = threadIdx.x >> LOG_WAVE;
  const int nw = (blockDim.x + WAVE - 1) >> LOG_WAVE;
  __syncthreads();
  if (lane_id() == 0) sh[w] = x;
  __syncthreads();
  T r = sh[0];
  if (threadIdx.x == 0)
    for (int i = 1; i < nw; ++i) r = r > sh[i] ? r : sh[i];
  return r;
}

// Exclusive scan over the workgroup (blockDim.x a multiple of 64, <= 1024).
// `sh` holds blockDim.x/64 ints. Returns the thread's exclusive prefix and
// the block total in `total` (all threads).
__device__ __forceinline__ int block_exclusive_scan(int x, int &total,
                                                    int *sh) {
  const int w = threadIdx.x >> LOG_WAVE;
  const int nw = blockDim.x >> LOG_WAVE;
  int wtot;
  int ex = wave_exclusive_scan(x, wtot);
  __syncthreads();
  if (lane_id() == 0) sh[w] = wtot;
  __syncthreads();
  int before = 0, tot = 0;
  for (int i = 0; i < nw; ++i) {
    int s = sh[i];
    before += (i < w) ? s : 0;
    tot += s;
  }
  total = tot;
  return ex + before;
}

// LDS writes of the wavefront visible to all its lanes (and earlier reads
// done) before the next access
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wavefront-staged append into a global array whose length lives at
// *counter: ids collect in an LDS buffer of CAP entries owned by the
// wavefront and reach the array in blocks (one atomic on the counter per
// flush). One atomic per call on a counter that every wavefront of the grid
// shares serialises at about 12 ns on MI355X; staging cuts the number of
// such atomics by up to CAP / 64. All 64 lanes call put() and flush()
// together; `cnt` is uniform across the wavefront. Entries past `cap` are
// dropped and raise *overflow.
template <int CAP>
struct wave_stage_t {
  int *buf;     // this wavefront's CAP-entry LDS buffer
  int cnt = 0;  // staged entries
  __device__ __forceinline__ void flush(int *counter, int *out, int64_t cap,
                                        int *overflow) {
    if (cnt == 0) return;
    wave_sync();  // every lane's buffer writes
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(counter, cnt);
    base = __shfl(base, 0, 64);
    for (int i = lane_id(); i < cnt; i += 64) {
      const int64_t slot = (int64_t)base + i;
      if (slot < cap)
        out[slot] = buf[i];
      else
        *overflow = 1;
    }
    wave_sync();  // buffer reads done before it is refilled
    cnt = 0;
  }
  // The workgroup's last flush: every wavefront's remaining entries go out
  // behind one atomic for the whole workgroup (a launch ends with one
  // atomic per workgroup instead of one per wavefront). Every thread of the
  // workgroup calls it; `s_cnt` is LDS of (waves + 1) ints.
  __device__ __forceinline__ void flush_block(int *counter, int *out, int64_t cap,
                                              int *overflow, int *s_cnt) {
    const int w = threadIdx.x >> LOG_WAVE, nw = blockDim.x >> LOG_WAVE;
    wave_sync();  // every lane's buffer writes
    if (lane_id() == 0) s_cnt[w] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int i = 0; i < nw; ++i) {
        const int c = s_cnt[i];
        s_cnt[i] = tot;
        tot += c;
      }
      s_cnt[nw] = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    const int64_t base = (int64_t)s_cnt[nw] + s_cnt[w];
    for (int i = lane_id(); i < cnt; i += 64) {
      const int64_t slot = base + i;
      if (slot < cap)
        out[slot] = buf[i];
      else
        *overflow = 1;
    }
    cnt = 0;
  }
  __device__ __forceinline__ void put(bool want, int u, int *counter, int *out,
                                      int64_t cap, int *overflow) {
    const uint64_t m = ballot(want);
    const int n = __popcll(m);
    if (n == 0) return;
    if (cnt + n > CAP) flush(counter, out, cap, overflow);
    if (want) buf[cnt + rank_below(m)] = u;
    cnt += n;
  }
};

// Index of the last element of sorted `a[0..len)` that is <= x, assuming
// a[0] <= x (owner lookup for edge-to-vertex mapping).
template <typename T>
__device__ __forceinline__ int owner_search(const T *a, int len, T x) {
  int lo = 0, hi = len;  // a[lo] <= x < a[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (a[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

}  // namespace gsw
