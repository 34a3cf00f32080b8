// Building the TWOD source slices on the device (layout: twod_t in
// data_structures/graph.h). Every in-adjacency row is sorted, so the part of
// row v that falls in slice p is the sub-range between two lower_bounds.
// Per slice: one pass computes each row's part (length and first edge), two
// scans turn the non-empty flags into row ids and the lengths into offsets,
// a scatter writes the compressed rows, and a tiled copy (one workgroup per
// 2048 output edges, owner rows found in LDS) fills the columns.
#pragma once
#include <cmath>
#include <cstdlib>
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

// part of row v with sources in [lo, hi): len[v], first[v] (absolute edge
// index of its first entry), nz[v] = (len > 0)
__global__ void __launch_bounds__(BLOCK)
    k_slice_len(const int64_t *rows, const int32_t *cols, int64_t n, int32_t lo,
                int32_t hi, int64_t *len, int64_t *first, int64_t *nz) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = rows[v], d = rows[v + 1] - s;
    const int64_t a = lower_bound_i32(cols + s, d, lo);
    const int64_t b = a + lower_bound_i32(cols + s + a, d - a, hi);
    len[v] = b - a;
    first[v] = s + a;
    nz[v] = (b > a) ? 1 : 0;
  }
}

// after the scans (off = exclusive scan of len, rid = of nz): the compressed
// rows of the slice; src[r] = first input edge of row r
__global__ void __launch_bounds__(BLOCK)
    k_slice_rows(int64_t n, const int64_t *off, const int64_t *rid,
                 const int64_t *first, int32_t *tgt, int64_t *soff,
                 int64_t *src) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = rid[v];
    if (rid[v + 1] > r) {
      tgt[r] = (int32_t)v;
      soff[r] = off[v];
      src[r] = first[v];
    }
    if (v == 0) soff[rid[n]] = off[n];
  }
}

// col[e] for the slice's edges, one workgroup per FILL_TILE outputs: rows
// have >= 1 edge, so at most FILL_TILE + 1 rows meet a tile; their offsets
// are staged in LDS and each output finds its row there
constexpr int FILL_TILE = 2048;
__global__ void __launch_bounds__(BLOCK)
    k_slice_fill(int64_t nr, int64_t ne, const int64_t *soff, const int64_t *src,
                 const int32_t *cols_in, int32_t *col) {
  __shared__ int64_t s_off[FILL_TILE + 2];
  __shared__ int64_t s_r0;
  for (int64_t e0 = (int64_t)blockIdx.x * FILL_TILE; e0 < ne;
       e0 += (int64_t)gridDim.x * FILL_TILE) {
    if (threadIdx.x == 0) s_r0 = last_le(soff, nr + 1, e0);
    __syncthreads();
    const int64_t r0 = s_r0;
    for (int j = threadIdx.x; j <= FILL_TILE + 1; j += BLOCK) {
      const int64_t r = r0 + j;
      s_off[j] = soff[r < nr ? r : nr];
    }
    __syncthreads();
    const int64_t e1 = MIN(ne, e0 + FILL_TILE);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += BLOCK) {
      const int k = owner_search(s_off, FILL_TILE + 2, e);
      col[e] = cols_in[src[r0 + k] + (e - s_off[k])];
    }
    __syncthreads();  // s_off / s_r0 are reused by the next tile
  }
}

// bin c of the rows (c = TWOD_BINS: the long rows): flag[r] = 1 (rows) or
// the chunk count (long rows), edges[r] = the row length if it is in c
__global__ void __launch_bounds__(BLOCK)
    k_bin_mark(const int64_t *off, int64_t nr, int c, int64_t *flag,
               int64_t *edges) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nr;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t len = off[r + 1] - off[r];
    int cls = TWOD_BINS;
    for (int k = TWOD_BINS - 1; k >= 0; --k)
      if (len <= twod_bin_len(k)) cls = k;
    const bool in = (cls == c);
    flag[r] = !in ? 0 : (c < TWOD_BINS ? 1 : (len + TWOD_CHUNK - 1) / TWOD_CHUNK);
    edges[r] = in ? len : 0;
  }
}

// after the scan of flag: list the rows of bin c (rows != nullptr) or the
// chunks of the long rows
__global__ void __launch_bounds__(BLOCK)
    k_bin_scatter(const int64_t *off, int64_t nr, const int64_t *pos,
                  int32_t *rows, int32_t *chunk_row, int64_t *chunk_first) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nr;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p0 = pos[r], cnt = pos[r + 1] - p0;
    if (cnt == 0) continue;
    if (rows) {
      rows[p0] = (int32_t)r;
    } else {
      for (int64_t j = 0; j < cnt; ++j) {
        chunk_row[p0 + j] = (int32_t)r;
        chunk_first[p0 + j] = off[r] + j * TWOD_CHUNK;
      }
    }
  }
}

// List the rows of slice `sl` by length and split a launch's workgroups
// over the parts; t1, t2 are scratch of sl.nr + 1 entries, bsum scan
// scratch.
inline void twod_bin(twod_slice_t &sl, int64_t *t1, int64_t *t2, int64_t *bsum) {
  int64_t edges[TWOD_BINS + 1] = {};
  const int grid = grid_for(sl.nr);
  for (int c = 0; c <= TWOD_BINS; ++c) {
    hipLaunchKernelGGL(k_bin_mark, dim3(grid), dim3(BLOCK), 0, stream(), sl.off,
                       sl.nr, c, t1, t2);
    H_ERR(hipGetLastError());
    scan_exclusive_inplace(t1, sl.nr, bsum);
    scan_exclusive_inplace(t2, sl.nr, bsum);
    int64_t cnt = 0;
    H_ERR(hipMemcpyAsync(&cnt, t1 + sl.nr, sizeof(int64_t), hipMemcpyDeviceToHost,
                         stream()));
    H_ERR(hipMemcpyAsync(&edges[c], t2 + sl.nr, sizeof(int64_t),
                         hipMemcpyDeviceToHost, stream()));
    sync();
    if (c < TWOD_BINS) {
      sl.bin_n[c] = cnt;
      H_ERR(hipMalloc((void **)&sl.bin_rows[c], sizeof(int32_t) * (cnt + 1)));
      hipLaunchKernelGGL(k_bin_scatter, dim3(grid), dim3(BLOCK), 0, stream(),
                         sl.off, sl.nr, t1, sl.bin_rows[c], nullptr, nullptr);
    } else {
      sl.n_chunks = cnt;
      H_ERR(hipMalloc((void **)&sl.chunk_row, sizeof(int32_t) * (cnt + 1)));
      H_ERR(hipMalloc((void **)&sl.chunk_first, sizeof(int64_t) * (cnt + 1)));
      hipLaunchKernelGGL(k_bin_scatter, dim3(grid), dim3(BLOCK), 0, stream(),
                         sl.off, sl.nr, t1, nullptr, sl.chunk_row, sl.chunk_first);
    }
    H_ERR(hipGetLastError());
    sync();
  }
  // workgroups per part, in proportion to its wavefront steps (a step is
  // 64 / group rows of a bin, or one chunk)
  double w[TWOD_BINS + 1], tot = 0;
  for (int c = 0; c < TWOD_BINS; ++c) {
    const int64_t per_step = WAVE / twod_bin_group(c);
    w[c] = (double)((sl.bin_n[c] + per_step - 1) / per_step);
  }
  w[TWOD_BINS] = (double)sl.n_chunks;
  for (int c = 0; c <= TWOD_BINS; ++c) tot += w[c];
  for (int c = 0; c <= TWOD_BINS; ++c) {
    int b = 0;
    if (w[c] > 0) {
      b = (int)std::ceil((double)MAX_GRID * w[c] / tot);
      // no more workgroups than the part has steps for their wavefronts
      const int64_t cap = ((int64_t)w[c] + BLOCK / WAVE - 1) / (BLOCK / WAVE);
      b = (int)MAX((int64_t)1, MIN((int64_t)b, cap));
    }
    sl.blocks[c] = b;
  }
  (void)edges;
}

// Build the slices of `g` once (no-op if built). `src_bytes` = bytes of
// source-side data one message reads (sets the slice width).
template <typename G> void twod_ensure(G &g, int64_t src_bytes) {
  GSW_REQUIRE(g.twod, E_INTERNAL_CHECK, "graph has no TWOD handle");
  twod_t &tw = *g.twod;
  if (tw.P) return;
  const int64_t n = g.nvertexs;
  GSW_REQUIRE(n >= 1 && src_bytes >= 1, E_INVALID, "bad TWOD build");
  // slice width: TWOD_SLICE_BYTES, or GSW_TWOD_SLICE_KB (a tuning knob)
  int64_t slice_bytes = TWOD_SLICE_BYTES;
  if (const char *e = std::getenv("GSW_TWOD_SLICE_KB")) {
    const long long kb = std::atoll(e);
    if (kb > 0) slice_bytes = (int64_t)kb << 10;
  }
  int64_t P = (n * src_bytes + slice_bytes - 1) / slice_bytes;
  P = MAX((int64_t)1, MIN((int64_t)TWOD_MAX, P));
  const int64_t width = (n + P - 1) / P;
  const int64_t *rows = g.dgr_row_ptr;
  const int32_t *cols = g.dgr_adj_list;
  int64_t *len = nullptr, *first = nullptr, *nz = nullptr, *bsum = nullptr;
  const int64_t nbs = (n + 1 + SCAN_TILE - 1) / SCAN_TILE + 2;
  H_ERR(hipMalloc((void **)&len, sizeof(int64_t) * (n + 1)));
  H_ERR(hipMalloc((void **)&first, sizeof(int64_t) * (n + 1)));
  H_ERR(hipMalloc((void **)&nz, sizeof(int64_t) * (n + 1)));
  H_ERR(hipMalloc((void **)&bsum, sizeof(int64_t) * nbs));
  tw.P = (int)P;
  for (int p = 0; p <= tw.P; ++p) tw.bound[p] = (int32_t)MIN(n, (int64_t)p * width);
  for (int p = 0; p < tw.P; ++p) {
    twod_slice_t &sl = tw.s[p];
    hipLaunchKernelGGL(k_slice_len, dim3(grid_for(n)), dim3(BLOCK), 0, stream(),
                       rows, cols, n, tw.bound[p], tw.bound[p + 1], len, first, nz);
    H_ERR(hipGetLastError());
    scan_exclusive_inplace(len, n, bsum);
    scan_exclusive_inplace(nz, n, bsum);
    H_ERR(hipMemcpyAsync(&sl.ne, len + n, sizeof(int64_t), hipMemcpyDeviceToHost,
                         stream()));
    H_ERR(hipMemcpyAsync(&sl.nr, nz + n, sizeof(int64_t), hipMemcpyDeviceToHost,
                         stream()));
    sync();
    int64_t *src = nullptr;
    H_ERR(hipMalloc((void **)&sl.tgt, sizeof(int32_t) * (sl.nr + 1)));
    H_ERR(hipMalloc((void **)&sl.off, sizeof(int64_t) * (sl.nr + 1)));
    H_ERR(hipMalloc((void **)&sl.col, sizeof(int32_t) * (sl.ne + 1)));
    H_ERR(hipMalloc((void **)&src, sizeof(int64_t) * (sl.nr + 1)));
    hipLaunchKernelGGL(k_slice_rows, dim3(grid_for(n)), dim3(BLOCK), 0, stream(),
                       n, len, nz, first, sl.tgt, sl.off, src);
    H_ERR(hipGetLastError());
    if (sl.ne > 0) {
      hipLaunchKernelGGL(k_slice_fill,
                         dim3(grid_for((sl.ne + FILL_TILE - 1) / FILL_TILE, 1, MAX_GRID)),
                         dim3(BLOCK), 0, stream(), sl.nr, sl.ne, sl.off, src, cols,
                         sl.col);
      H_ERR(hipGetLastError());
    }
    sync();
    hipFree(src);
    if (sl.nr > 0) twod_bin(sl, len, nz, bsum);
  }
  hipFree(len);
  hipFree(first);
  hipFree(nz);
  hipFree(bsum);
}

}  // namespace gsw
