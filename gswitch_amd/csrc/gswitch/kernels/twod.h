// Building the TWOD source slices on the device (layout: twod_t in
// data_structures/graph.h). Slice boundaries split the source ids at equal
// out-edge counts (one eighth of the edges each, so the eight XCDs get equal
// work); each slice is a full CSR over all targets whose rows hold only the
// in-edges from that slice's sources. Every adjacency row is sorted, so a
// row's part in a slice is the sub-range between two lower_bounds.
#pragma once
#include <vector>

#include "../data_structures/active_set.h"
#include "../data_structures/graph.h"
#include "scan.h"

namespace gsw {

__device__ __forceinline__ int64_t lower_bound_i32(const int32_t *a, int64_t len,
                                                   int32_t x) {
  int64_t lo = 0, hi = len;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// rp[p][v] <- number of in-edges of v whose source lies in slice p
__global__ void __launch_bounds__(BLOCK)
    k_twod_len(const int64_t *rows, const int32_t *cols, int64_t n,
               twod_t tw) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = rows[v], len = rows[v + 1] - s;
    int64_t a = lower_bound_i32(cols + s, len, tw.bound[0]);
    for (int p = 0; p < tw.P; ++p) {
      const int64_t b = lower_bound_i32(cols + s, len, tw.bound[p + 1]);
      tw.rp[p][v] = b - a;
      a = b;
    }
  }
}

__global__ void __launch_bounds__(BLOCK)
    k_twod_fill(const int64_t *rows, const int32_t *cols, int64_t n,
                twod_t tw) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = rows[v], len = rows[v + 1] - s;
    int64_t a = lower_bound_i32(cols + s, len, tw.bound[0]);
    for (int p = 0; p < tw.P; ++p) {
      const int64_t b = lower_bound_i32(cols + s, len, tw.bound[p + 1]);
      int32_t *dst = tw.col[p] + tw.rp[p][v];
      for (int64_t j = a; j < b; ++j) dst[j - a] = cols[s + j];
      a = b;
    }
  }
}

// Build the slices of `g` once (no-op if built). `bsum` is scan scratch of
// ceil(n / SCAN_TILE) + 2 entries.
template <typename G> void twod_ensure(G &g, int P, int64_t *bsum) {
  GSW_REQUIRE(g.twod, E_INTERNAL_CHECK, "graph has no TWOD handle");
  twod_t &tw = *g.twod;
  if (tw.P) return;
  GSW_REQUIRE(P >= 1 && P <= TWOD_MAX, E_INVALID, "bad TWOD slice count");
  const int64_t n = g.nvertexs;
  // source histogram = forward out-degrees; boundaries at equal edge prefix
  std::vector<int64_t> rp((size_t)n + 1);
  H_ERR(hipMemcpy(rp.data(), g.dg_row_ptr, sizeof(int64_t) * (n + 1),
                  hipMemcpyDeviceToHost));
  const int64_t m = rp[n];
  tw.bound[0] = 0;
  for (int p = 1; p < P; ++p) {
    const int64_t target = m * p / P;
    int64_t v = std::lower_bound(rp.begin(), rp.end(), target) - rp.begin();
    v = MIN(v, n);
    tw.bound[p] = (int32_t)MAX((int64_t)tw.bound[p - 1], v);
  }
  tw.bound[P] = (int32_t)n;
  for (int p = 0; p < P; ++p)
    H_ERR(hipMalloc((void **)&tw.rp[p], sizeof(int64_t) * (n + 1)));
  tw.P = P;
  hipLaunchKernelGGL(k_twod_len, dim3(grid_for(n)), dim3(BLOCK), 0, stream(),
                     (const int64_t *)g.dgr_row_ptr,
                     (const int32_t *)g.dgr_adj_list, n, tw);
  H_ERR(hipGetLastError());
  std::vector<int64_t> tot(P);
  for (int p = 0; p < P; ++p) {
    scan_exclusive_inplace(tw.rp[p], n, bsum);
    H_ERR(hipMemcpyAsync(&tot[p], tw.rp[p] + n, sizeof(int64_t),
                         hipMemcpyDeviceToHost, stream()));
  }
  sync();
  for (int p = 0; p < P; ++p)
    H_ERR(hipMalloc((void **)&tw.col[p], sizeof(int32_t) * (tot[p] + 1)));
  hipLaunchKernelGGL(k_twod_fill, dim3(grid_for(n)), dim3(BLOCK), 0, stream(),
                     (const int64_t *)g.dgr_row_ptr,
                     (const int32_t *)g.dgr_adj_list, n, tw);
  H_ERR(hipGetLastError());
  sync();
}

}  // namespace gsw
