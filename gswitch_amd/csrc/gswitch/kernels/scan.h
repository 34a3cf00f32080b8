// Device-wide exclusive scan over int64, and the upper-bound search used by
// the edge-balanced (STRICT) partition. These replace moderngpu's
// mgpu::scan<scan_type_exc> and mgpu::sorted_search<bounds_lower>
// (called at src/kernel_libs/expand_VC_STRICT.cuh:540-552).
//
// scan_exclusive_inplace(a, N): a[i] <- sum(a[0..i)), a[N] <- total.
// Three launches: per-tile sums, one-workgroup scan of the tile sums,
// per-tile scan with the tile offset. Tiles are 256 threads x 8 items.
#pragma once
#include "../utils/runtime.h"
#include "../utils/wave.h"

namespace gsw {

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;  // 2048

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan_any(T x) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x;
}

// exclusive scan across the workgroup for 64-bit values; sh holds nw+1 T
template <typename T>
__device__ __forceinline__ T block_exclusive_scan_any(T x, T &total, T *sh) {
  const int w = threadIdx.x >> LOG_WAVE;
  const int nw = blockDim.x >> LOG_WAVE;
  T inc = wave_inclusive_scan_any(x);
  __syncthreads();
  if (lane_id() == 63) sh[w] = inc;
  __syncthreads();
  T before = 0, tot = 0;
  for (int i = 0; i < nw; ++i) {
    T s = sh[i];
    before += (i < w) ? s : (T)0;
    tot += s;
  }
  total = tot;
  return inc - x + before;
}

__global__ void __launch_bounds__(BLOCK)
    k_scan_tile_sums(const int64_t *__restrict__ a, int64_t N,
                     int64_t *__restrict__ bsum) {
  __shared__ int64_t sh[BLOCK / WAVE];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + (int64_t)k * BLOCK + threadIdx.x;
    if (i < N) s += a[i];
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) bsum[blockIdx.x] = s;
}

// one workgroup of 1024 threads scans nb tile sums; writes total to both
// bsum[nb] and a[N]
__global__ void __launch_bounds__(1024)
    k_scan_tile_offsets(int64_t *__restrict__ bsum, int nb,
                        int64_t *__restrict__ a, int64_t N) {
  __shared__ int64_t sh[1024 / WAVE + 1];
  const int per = (nb + 1023) / 1024;
  const int lo = threadIdx.x * per;
  int64_t local = 0;
  for (int i = lo; i < lo + per && i < nb; ++i) local += bsum[i];
  int64_t total;
  int64_t run = block_exclusive_scan_any(local, total, sh);
  for (int i = lo; i < lo + per && i < nb; ++i) {
    int64_t x = bsum[i];
    bsum[i] = run;
    run += x;
  }
  if (threadIdx.x == 0) {
    bsum[nb] = total;
    a[N] = total;
  }
}

__global__ void __launch_bounds__(BLOCK)
    k_scan_tile_apply(int64_t *__restrict__ a, int64_t N,
                      const int64_t *__restrict__ bsum) {
  __shared__ int64_t sh[BLOCK / WAVE + 1];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE +
                       (int64_t)threadIdx.x * SCAN_ITEMS;
  int64_t x[SCAN_ITEMS];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + k;
    x[k] = (i < N) ? a[i] : 0;
    s += x[k];
  }
  int64_t total;
  int64_t run = block_exclusive_scan_any(s, total, sh) + bsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + k;
    if (i < N) a[i] = run;
    run += x[k];
  }
}

// Host driver. `bsum` must hold ceil(N / 2048) + 1 entries.
inline void scan_exclusive_inplace(int64_t *a, int64_t N, int64_t *bsum) {
  const int64_t nb64 = (N + SCAN_TILE - 1) / SCAN_TILE;
  GSW_REQUIRE(nb64 < (1 << 26), E_INVALID, "scan too large");
  const int nb = (int)(nb64 > 0 ? nb64 : 0);
  hipStream_t s = stream();
  if (nb > 0)
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(nb), dim3(BLOCK), 0, s, a, N, bsum);
  hipLaunchKernelGGL(k_scan_tile_offsets, dim3(1), dim3(1024), 0, s, bsum, nb, a, N);
  if (nb > 0)
    hipLaunchKernelGGL(k_scan_tile_apply, dim3(nb), dim3(BLOCK), 0, s, a, N, bsum);
  H_ERR(hipGetLastError());
}

// Largest k in [0, len) with a[k] <= x (a sorted ascending, a[0] <= x).
__device__ __forceinline__ int64_t last_le(const int64_t *a, int64_t len,
                                           int64_t x) {
  int64_t lo = 0, hi = len;
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) >> 1;
    if (a[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

}  // namespace gsw
