// Frontier generation: bitmap -> queue.
//
// The reference's live path (__filter_active / __filter_inactive /
// __filter_unfixed, src/kernel_libs/filter.cuh:30-112) walks byte words with
// a 32-lane scan and one atomic per warp, giving a queue sorted only inside
// each warp's stripe. Here:
//  * unsorted: each lane owns one 64-bit word; a DPP wave scan of popcounts
//    and one atomicAdd per wavefront reserve the output range;
//  * sorted:   workgroups own contiguous word ranges; pass 1 counts per
//    workgroup, pass 2 offsets each workgroup by the counts before it and
//    writes ids in ascending order, so the queue is globally sorted.
#pragma once
#include "../data_structures/active_set.h"
#include "../utils/wave.h"

namespace gsw {

// word w of the target set: Active / Inactive / All (= not Fixed), tail
// bits beyond n cleared
__device__ __forceinline__ uint64_t target_word(const active_set_t &as,
                                                int64_t w, Status target,
                                                int64_t n) {
  const uint64_t a = as.bitmap.active.dg_bits[w];
  const uint64_t i = as.bitmap.inactive.dg_bits[w];
  uint64_t m = (target == Active) ? a : (target == Inactive) ? ~i : (a | ~i);
  const int64_t rem = n - (w << 6);
  if (rem < 64) m &= (rem <= 0) ? 0ull : ((1ull << rem) - 1);
  return m;
}

__device__ __forceinline__ void write_word_ids(uint64_t m, int64_t w,
                                               int32_t *out, int64_t pos) {
  while (m) {
    int b = __builtin_ctzll(m);
    out[pos++] = (int32_t)((w << 6) + b);
    m &= m - 1;
  }
}

// Each thread owns FILTER_WPT consecutive words; a workgroup scan of their
// popcounts and one atomicAdd per workgroup step reserve the output range
// (one atomic per 64 Ki vertices: a counter shared by every wavefront of
// the grid serialises its atomics).
constexpr int FILTER_WPT = 4;
__global__ void __launch_bounds__(BLOCK)
    k_filter_unsorted(active_set_t as, Status target, int64_t n,
                      int32_t *out, int32_t *out_size) {
  __shared__ int sh[BLOCK / WAVE];
  __shared__ int s_base;
  const int64_t nw = (n + 63) >> 6;
  const int64_t step = (int64_t)gridDim.x * BLOCK * FILTER_WPT;
  for (int64_t w0 = (int64_t)blockIdx.x * BLOCK * FILTER_WPT; w0 < nw; w0 += step) {
    const int64_t w = w0 + (int64_t)threadIdx.x * FILTER_WPT;
    uint64_t m[FILTER_WPT];
    int c = 0;
#pragma unroll
    for (int k = 0; k < FILTER_WPT; ++k) {
      m[k] = (w + k < nw) ? target_word(as, w + k, target, n) : 0ull;
      c += __popcll(m[k]);
    }
    int tot;
    int pos = block_exclusive_scan(c, tot, sh);
    if (threadIdx.x == 0) s_base = tot ? atomicAdd(out_size, tot) : 0;
    __syncthreads();
    pos += s_base;
#pragma unroll
    for (int k = 0; k < FILTER_WPT; ++k) {
      write_word_ids(m[k], w + k, out, pos);
      pos += __popcll(m[k]);
    }
    __syncthreads();  // s_base and sh are reused by the next step
  }
}

// pass 1: per-workgroup counts over contiguous word ranges
__global__ void __launch_bounds__(BLOCK)
    k_filter_sorted_count(active_set_t as, Status target, int64_t n,
                          int64_t words_per_block, int32_t *block_count) {
  __shared__ int sh[BLOCK / WAVE];
  const int64_t nw = (n + 63) >> 6;
  const int64_t lo = (int64_t)blockIdx.x * words_per_block;
  const int64_t hi = MIN(nw, lo + words_per_block);
  int c = 0;
  for (int64_t w = lo + threadIdx.x; w < hi; w += blockDim.x)
    c += __popcll(target_word(as, w, target, n));
  c = block_sum(c, sh);
  if (threadIdx.x == 0) block_count[blockIdx.x] = c;
}

// pass 2: ordered write
__global__ void __launch_bounds__(BLOCK)
    k_filter_sorted_write(active_set_t as, Status target, int64_t n,
                          int64_t words_per_block, const int32_t *block_count,
                          int32_t *out, int32_t *out_size) {
  __shared__ int sh[BLOCK / WAVE];
  const int64_t nw = (n + 63) >> 6;
  int before = 0;
  for (int i = threadIdx.x; i < (int)blockIdx.x; i += blockDim.x)
    before += block_count[i];
  before = block_sum(before, sh);
  __shared__ int s_base;
  if (threadIdx.x == 0) s_base = before;
  __syncthreads();
  int base = s_base;
  const int64_t lo = (int64_t)blockIdx.x * words_per_block;
  const int64_t hi = MIN(nw, lo + words_per_block);
  for (int64_t w0 = lo; w0 < hi; w0 += blockDim.x) {
    const int64_t w = w0 + threadIdx.x;
    const uint64_t m = (w < hi) ? target_word(as, w, target, n) : 0ull;
    int tot;
    const int ex = block_exclusive_scan(__popcll(m), tot, sh);
    write_word_ids(m, w, out, base + ex);
    base += tot;
    __syncthreads();
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *out_size = base;
}

}  // namespace gsw
