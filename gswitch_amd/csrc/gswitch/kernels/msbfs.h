// 64-source bitmask BFS (multi-source BFS with one uint64 per vertex).
//
// No counterpart kernel in the reference (application/mbfs.cu runs the same
// single-source BFS three times); this is the batched form of its BFS: bit s
// of visit[v] is set once v is reachable from source s, bit s of front[v]
// once v was first reached from s at the previous level. A level is either
//  pull: every vertex with unvisited source bits ORs its neighbours' front
//        masks (stopping once all its open bits are covered); vertices with
//        more than MS_WIDE edges are scanned by the whole wavefront;
//  push: every vertex with a non-empty front ORs it into its neighbours'
//        next masks (64-bit atomicOr), then a sweep keeps the new bits.
// Per level the kernels also count, per source, how many vertices were newly
// reached (64 ballot/popcount pairs per wavefront), from which the host
// builds a per-source sum of levels -- a checksum of all 64 BFS trees.
#pragma once
#include "../utils/runtime.h"
#include "../utils/wave.h"

namespace gsw {

constexpr int MS_WIDE = 256;  // degree above which the wavefront cooperates

struct ms_counters_t {
  unsigned long long newly;         // vertices with new bits this level
  unsigned long long newly_deg;     // their degree sum
  unsigned long long per_src[64];   // newly reached vertices per source
  unsigned long long pad[14];       // one slot = five 128-byte lines
};
// Workgroups add their counts into one of MS_SLOTS partial copies (slot
// blockIdx.x % MS_SLOTS); the host sums the copies. Atomics that every
// workgroup of the grid sends to the same lines serialise on MI355X.
constexpr int MS_SLOTS = 64;
static_assert(sizeof(ms_counters_t) == 640, "slot size");

// per-source counts of the 64 lanes' `bits`; lane s accumulates source s
__device__ __forceinline__ void ms_count_sources(uint64_t bits,
                                                 unsigned long long &cnt) {
  const int lane = lane_id();
#pragma unroll 8
  for (int s = 0; s < 64; ++s) {
    const uint64_t b = ballot((bits >> s) & 1ull);
    if (lane == s) cnt += __popcll(b);
  }
}

// Workgroup reduction of the counts, then one set of atomics into the
// workgroup's slot. Every thread of the workgroup calls it.
__device__ __forceinline__ void ms_flush(ms_counters_t *c,
                                         unsigned long long newly,
                                         unsigned long long ndeg,
                                         unsigned long long cnt) {
  constexpr int NW = BLOCK / WAVE;
  __shared__ unsigned long long s_src[NW][WAVE];
  __shared__ unsigned long long s_tot[NW][2];
  const int w = threadIdx.x >> LOG_WAVE, lane = lane_id();
  newly = wave_sum_any(newly);
  ndeg = wave_sum_any(ndeg);
  s_src[w][lane] = cnt;
  if (lane == 0) {
    s_tot[w][0] = newly;
    s_tot[w][1] = ndeg;
  }
  __syncthreads();
  if (w == 0) {
    ms_counters_t *slot = c + (blockIdx.x & (MS_SLOTS - 1));
    unsigned long long t = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) t += s_src[i][lane];
    if (t) atomicAdd(&slot->per_src[lane], t);
    if (lane < 2) {
      unsigned long long x = 0;
#pragma unroll
      for (int i = 0; i < NW; ++i) x += s_tot[i][lane];
      if (x) atomicAdd(lane == 0 ? &slot->newly : &slot->newly_deg, x);
    }
  }
}

// Bottom-up level over the in-edges (rows/cols: the reverse CSR of a
// directed graph). Writes next[v] = newly reached bits for every v and
// counts the out-degrees (out_rows) of the newly reached vertices, the
// next level's push work.
__global__ void __launch_bounds__(BLOCK)
    k_msbfs_pull(const int64_t *__restrict__ rows,
                 const int32_t *__restrict__ cols,
                 const int64_t *__restrict__ out_rows, int64_t n, uint64_t all,
                 uint64_t *__restrict__ visit,
                 const uint64_t *__restrict__ front,
                 uint64_t *__restrict__ next, ms_counters_t *c) {
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  unsigned long long newly = 0, ndeg = 0, cnt = 0;
  for (int64_t base = wave << 6; base < n; base += nwaves << 6) {
    const int64_t v = base + lane;
    uint64_t vis = 0, acc = 0;
    int64_t s = 0, e = 0;
    if (v < n) {
      vis = visit[v];
      if (vis != all) {
        s = rows[v];
        e = rows[v + 1];
      }
    }
    const uint64_t open = all & ~vis;
    const bool wide = (e - s) > MS_WIDE;
    if (!wide) {
      for (int64_t j = s; j < e; ++j) {
        acc |= front[cols[j]];
        if ((acc & open) == open) break;
      }
    }
    // wavefront-cooperative scan of the wide vertices, one at a time
    uint64_t wmask = ballot(wide);
    while (wmask) {
      const int w = __builtin_ctzll(wmask);
      wmask &= wmask - 1;
      const int64_t ws = __shfl(s, w, 64), we = __shfl(e, w, 64);
      const uint64_t wopen = __shfl(open, w, 64);
      uint64_t wacc = 0;
      for (int64_t j0 = ws; j0 < we; j0 += WAVE) {
        const int64_t j = j0 + lane;
        uint64_t x = (j < we) ? front[cols[j]] : 0ull;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o, 64);
        wacc |= x;
        if ((wacc & wopen) == wopen) break;
      }
      if (lane == w) acc = wacc;
    }
    const uint64_t nw = acc & open;
    if (v < n) {
      next[v] = nw;
      if (nw) {
        visit[v] = vis | nw;
        ++newly;
        ndeg += (uint64_t)(out_rows[v + 1] - out_rows[v]);
      }
    }
    ms_count_sources(nw, cnt);
  }
  ms_flush(c, newly, ndeg, cnt);
}

// Top-down level, scatter part: next[u] |= front[v] & ~visit[u].
__global__ void __launch_bounds__(BLOCK)
    k_msbfs_push(const int64_t *__restrict__ rows,
                 const int32_t *__restrict__ cols, int64_t n,
                 const uint64_t *__restrict__ visit,
                 const uint64_t *__restrict__ front, uint64_t *next) {
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  for (int64_t base = wave << 6; base < n; base += nwaves << 6) {
    const int64_t v = base + lane;
    const uint64_t f = (v < n) ? front[v] : 0ull;
    int64_t s = 0, e = 0;
    if (f) {
      s = rows[v];
      e = rows[v + 1];
    }
    const bool wide = (e - s) > MS_WIDE;
    if (!wide)
      for (int64_t j = s; j < e; ++j) {
        const int32_t u = cols[j];
        const uint64_t b = f & ~visit[u];
        if (b && (next[u] & b) != b)
          atomicOr((unsigned long long *)&next[u], (unsigned long long)b);
      }
    uint64_t wmask = ballot(wide);
    while (wmask) {
      const int w = __builtin_ctzll(wmask);
      wmask &= wmask - 1;
      const int64_t ws = __shfl(s, w, 64), we = __shfl(e, w, 64);
      const uint64_t wf = __shfl(f, w, 64);
      for (int64_t j = ws + lane; j < we; j += WAVE) {
        const int32_t u = cols[j];
        const uint64_t b = wf & ~visit[u];
        if (b && (next[u] & b) != b)
          atomicOr((unsigned long long *)&next[u], (unsigned long long)b);
      }
    }
  }
}

// First top-down level: the frontier is the sources, which are often the
// graph's hubs, so their edges are spread evenly over the whole grid (edge k
// of the sources' concatenated lists, found by binary search over their
// degree prefix) instead of one wavefront per source vertex.
struct ms_seeds_t {
  int nsrc;
  int32_t src[64];
  int64_t pre[65];  // exclusive prefix of the sources' degrees
};

__global__ void __launch_bounds__(BLOCK)
    k_msbfs_push_seeds(const int64_t *__restrict__ rows,
                       const int32_t *__restrict__ cols,
                       const uint64_t *__restrict__ visit,
                       const uint64_t *__restrict__ front, uint64_t *next,
                       ms_seeds_t sd) {
  __shared__ int64_t s_pre[65];
  __shared__ int64_t s_row[64];
  __shared__ uint64_t s_bits[64];
  const int t = threadIdx.x;
  if (t <= sd.nsrc) s_pre[t] = sd.pre[t];
  if (t < sd.nsrc) {
    s_row[t] = rows[sd.src[t]];
    s_bits[t] = front[sd.src[t]];
  }
  __syncthreads();
  const int64_t total = s_pre[sd.nsrc];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + t; k < total; k += stride) {
    int lo = 0, hi = sd.nsrc;  // the last source with s_pre[lo] <= k
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (s_pre[mid] <= k)
        lo = mid;
      else
        hi = mid;
    }
    const int32_t u = cols[s_row[lo] + (k - s_pre[lo])];
    const uint64_t b = s_bits[lo] & ~visit[u];
    if (b && (next[u] & b) != b)
      atomicOr((unsigned long long *)&next[u], (unsigned long long)b);
  }
}

// Top-down level, sweep part: keep only new bits, fold them into visit.
__global__ void __launch_bounds__(BLOCK)
    k_msbfs_sweep(const int64_t *__restrict__ rows, int64_t n,
                  uint64_t *__restrict__ visit, uint64_t *__restrict__ next,
                  ms_counters_t *c) {
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  unsigned long long newly = 0, ndeg = 0, cnt = 0;
  for (int64_t base = wave << 6; base < n; base += nwaves << 6) {
    const int64_t v = base + lane_id();
    uint64_t nw = 0;
    if (v < n) {
      const uint64_t x = next[v];
      if (x) {
        const uint64_t vis = visit[v];
        nw = x & ~vis;
        next[v] = nw;
        if (nw) {
          visit[v] = vis | nw;
          ++newly;
          ndeg += (uint64_t)(rows[v + 1] - rows[v]);
        }
      }
    }
    ms_count_sources(nw, cnt);
  }
  ms_flush(c, newly, ndeg, cnt);
}

__global__ void __launch_bounds__(BLOCK)
    k_msbfs_seed(uint64_t *visit, uint64_t *front, const int32_t *src,
                 int nsrc) {
  const int s = threadIdx.x;
  if (s < nsrc) {
    atomicOr((unsigned long long *)&visit[src[s]], 1ull << s);
    atomicOr((unsigned long long *)&front[src[s]], 1ull << s);
  }
}

}  // namespace gsw
