// Connected components by tree hooking over the CSR, with subgraph sampling
// (Afforest: Sutton, Ben-Nun, Bar-Noy, "Optimizing Parallel Graph
// Connectivity Computation via Subgraph Sampling", IPDPS 2018).
//
// p[v] is a parent pointer with p[v] <= v; roots have p[r] = r. link(u, v)
// hooks the larger of the two roots under the smaller one with a
// compare-and-swap, so every tree's root is the smallest id in it and, once a
// component is one tree, its root is the component's minimum vertex id: the
// canonical label of the reference's CC (application/cc.cu:34-66). link
// never waits on another lane: each failed CAS means the root it tried moved
// to a smaller value, so the loop ends.
//
// Phases (a fixed sequence of launches, no host decisions in between for
// the batched form, so it can live in a captured graph):
//  1. p = identity;
//  2. sampling: every vertex links its first neighbour, compress, links its
//     second neighbour, compress;
//  3. skip set S: vertices whose root is a chosen root -- the most frequent
//     root of a vertex sample for one graph, or each instance's first vertex
//     (its smallest id) for a disjoint batch. Vertices outside S link all
//     their remaining edges; a vertex in S skips the edges that stay inside
//     its instance (their other end is either in the same tree already or
//     outside S, and then links the edge itself) and links only edges that
//     leave its instance, so a batch whose instances share edges stays
//     correct;
//  4. compress: p[v] = root(v).
// Undirected graphs only use S (an edge must be visible from both ends);
// a directed graph links every out-edge of every vertex (weak components).
#pragma once
#include "../utils/runtime.h"
#include "../utils/wave.h"

namespace gsw {

constexpr int CC_WIDE = 128;  // degree above which a wavefront shares a row

__device__ __forceinline__ int cc_get(const int *p, int64_t i) {
  return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// join the trees of u and v: hook the larger root under the smaller one
// with a CAS, walking up from the parents (two steps at a time on the
// larger side) while neither is a root. Measured against finds that halve
// the path as they walk (Jaiganesh and Burtscher, HPDC 2018): on a union of
// 5000 random graphs the halving stores cost 180 ms against 74 ms for the
// whole batched BFS + CC, and the compress after each sampling round
// already keeps grid chains short (4096-square grid 3.3 vs 2.7 ms).
__device__ __forceinline__ void cc_link(int u, int v, int *p) {
  int a = cc_get(p, u), b = cc_get(p, v);
  while (a != b) {
    const int hi = a > b ? a : b, lo = a > b ? b : a;
    const int ph = cc_get(p, hi);
    if (ph == lo) return;
    if (ph == hi && atomicCAS(p + hi, hi, lo) == hi) return;
    a = cc_get(p, cc_get(p, hi));
    b = cc_get(p, lo);
  }
}

__global__ void __launch_bounds__(BLOCK) k_cc_init(int *p, int64_t n) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    p[v] = (int)v;
}

// link every vertex with its r-th neighbour (with `skip_lv`, only the
// vertices whose skip_lv entry is negative)
__global__ void __launch_bounds__(BLOCK)
    k_cc_sample(const int64_t *__restrict__ rows, const int32_t *__restrict__ cols,
                int64_t n, int r, int *p, const int *__restrict__ skip_lv) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (skip_lv && skip_lv[v] >= 0) continue;
    const int64_t s = rows[v];
    if (rows[v + 1] - s > r) cc_link((int)v, cols[s + r], p);
  }
}

// p[v] <- root of v (no links run meanwhile, so p is a forest whose roots
// stay put). Every vertex moves its own pointer to its grandparent until
// the parent is a root; the pointers it reads are being shortened by their
// own vertices at the same time, so a chain of length d collapses in about
// log d steps per vertex instead of each vertex walking it alone (sampled
// links on a grid make chains of thousands).
__global__ void __launch_bounds__(BLOCK) k_cc_compress(int *p, int64_t n) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    int x = cc_get(p, v);
    int y = cc_get(p, x);
    while (y != x) {
      __hip_atomic_store(p + v, y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      x = y;
      y = cc_get(p, x);
    }
  }
}

// ---- choosing the skipped component on the device ------------------------
// A fixed sample of vertex ids (SplitMix64 output mix over key + (i+1)*phi,
// scaled to [0, n) by a multiply-high: the host generators' counter RNG),
// then the most frequent root among the sampled vertices. No host read, so
// the whole CC is one asynchronous launch sequence.
constexpr int CC_MAX_SAMPLES = 4096;
constexpr int CC_MODE_BLOCK = 1024;
constexpr uint64_t CC_SAMPLE_KEY = 0xCC5A3D1Eull;

__device__ __forceinline__ uint64_t cc_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(BLOCK)
    k_cc_sample_ids(int32_t *ids, int k, int64_t n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const uint64_t key = cc_mix64(cc_mix64(CC_SAMPLE_KEY));
  const uint64_t r = cc_mix64(key + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull);
  ids[i] = (int32_t)__umul64hi(r, (uint64_t)n);
}

// *skip <- the most frequent value of p[ids[0..k)] (ties: the smallest).
// One workgroup of CC_MODE_BLOCK threads, 1 <= k <= CC_MAX_SAMPLES: bitonic
// sort in LDS, then each run start measures its run and a 64-bit max of
// (run length, INT_MAX - value) picks the winner.
__global__ void __launch_bounds__(CC_MODE_BLOCK)
    k_cc_mode(const int *p, const int32_t *ids, int k, int *skip) {
  __shared__ int s[CC_MAX_SAMPLES];
  __shared__ unsigned long long best;
  int k2 = 1;
  while (k2 < k) k2 <<= 1;
  for (int i = threadIdx.x; i < k2; i += blockDim.x)
    s[i] = i < k ? cc_get(p, ids[i]) : INT_MAX;
  if (threadIdx.x == 0) best = 0;
  for (int size = 2; size <= k2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      for (int t = threadIdx.x; t < (k2 >> 1); t += blockDim.x) {
        const int i = 2 * t - (t & (stride - 1)), j = i + stride;
        const int a = s[i], b = s[j];
        if ((a > b) == ((i & size) == 0)) {
          s[i] = b;
          s[j] = a;
        }
      }
    }
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    if (i > 0 && s[i] == s[i - 1]) continue;
    int j = i + 1;
    while (j < k && s[j] == s[i]) ++j;
    const unsigned long long key =
        ((unsigned long long)(j - i) << 32) | (unsigned)(INT_MAX - s[i]);
    atomicMax(&best, key);
  }
  __syncthreads();
  if (threadIdx.x == 0) *skip = INT_MAX - (int)(unsigned)(best & 0xFFFFFFFFull);
}

// Instance of v: the largest i with offsets[i] <= v (offsets ascending,
// offsets[0] = 0, count + 1 entries).
__device__ __forceinline__ int64_t cc_instance(const int64_t *off, int64_t count,
                                               int64_t v) {
  int64_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (off[mid] <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Instance of vertex v of a wavefront whose lanes hold base .. base + 63:
// one search for base (every lane the same addresses), then forward steps.
__device__ __forceinline__ int64_t cc_instance_wave(const int64_t *off, int64_t count,
                                                    int64_t base, int64_t v) {
  int64_t i = cc_instance(off, count, base);
  while (i + 1 < count && off[i + 1] <= v) ++i;
  return i;
}

// For a batch of instances whose edges stay inside their instance (checked
// by k_cc_cross) and BFS levels from every instance's first vertex: the
// vertices the BFS reached form the component of that first vertex, the
// smallest id of the instance, so that id is their label; every other
// vertex starts as its own root.
__global__ void __launch_bounds__(BLOCK)
    k_cc_seed_levels(int *p, const int *__restrict__ lv, const int64_t *__restrict__ off,
                     int64_t count, int64_t n) {
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  for (int64_t base = wave << 6; base < n; base += nwaves << 6) {
    const int64_t v = base + lane;
    if (v < n)
      p[v] = lv[v] >= 0 ? (int)off[cc_instance_wave(off, count, base, v)] : (int)v;
  }
}

// *flag = 1 if an edge leaves its source's instance.
__global__ void __launch_bounds__(BLOCK)
    k_cc_cross(const int64_t *__restrict__ rows, const int32_t *__restrict__ cols,
               int64_t n, const int64_t *__restrict__ off, int64_t count, int *flag) {
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  bool out = false;
  for (int64_t base = wave << 6; base < n; base += nwaves << 6) {
    const int64_t v = base + lane;
    int64_t s = 0, e = 0, lo = 0, hi = 0;
    if (v < n) {
      s = rows[v];
      e = rows[v + 1];
      const int64_t i = cc_instance_wave(off, count, base, v);
      lo = off[i];
      hi = off[i + 1];
    }
    // columns of a row are sorted: its first and last entries decide
    if (e > s && (cols[s] < lo || cols[e - 1] >= hi)) out = true;
  }
  if (ballot(out) && lane == 0) atomicOr(flag, 1);
}

// Phase 3. skip_root >= 0: S = {v : p[v] == skip_root} (one graph);
// skip_root == -1: S = {v : p[v] == first vertex of v's instance} with
// `off` the count + 1 instance offsets; skip_root == -2: no skipping. A
// non-null skip_dev replaces skip_root by *skip_dev (written by k_cc_mode).
// Edges [from, deg) of each row are visited (the sampled ones are done).
// One wavefront per 64 vertices; rows longer than CC_WIDE are walked by
// the whole wavefront, one at a time. `disjoint`: no edge leaves an
// instance (skip_root == -1), so a vertex of S links nothing.
__global__ void __launch_bounds__(BLOCK)
    k_cc_rest(const int64_t *__restrict__ rows, const int32_t *__restrict__ cols,
              int64_t n, int from, int skip_arg, const int *skip_dev,
              const int64_t *__restrict__ off, int64_t count, int *p,
              int disjoint) {
  const int skip_root = skip_dev ? *skip_dev : skip_arg;
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  for (int64_t base = wave << 6; base < n; base += nwaves << 6) {
    const int64_t v = base + lane;
    int64_t s = 0, e = 0, lo = 0, hi = 0;
    bool in_s = false;
    if (v < n) {
      s = rows[v] + from;
      e = rows[v + 1];
      if (s > e) s = e;
      const int root = cc_get(p, v);
      if (skip_root >= 0) {
        in_s = root == skip_root;
        hi = n;
      } else if (skip_root == -1) {
        const int64_t i = cc_instance(off, count, v);
        lo = off[i];
        hi = off[i + 1];
        in_s = root == (int)lo;
      }
      // one graph, or instances with no edge between them: a vertex of S
      // has nothing to link
      if (in_s && (skip_root >= 0 || disjoint)) e = s;
    }
    const bool wide = (e - s) > CC_WIDE;
    if (!wide)
      for (int64_t j = s; j < e; ++j) {
        const int u = cols[j];
        if (!in_s || u < lo || u >= hi) cc_link((int)v, u, p);
      }
    uint64_t wmask = ballot(wide);
    while (wmask) {
      const int w = __builtin_ctzll(wmask);
      wmask &= wmask - 1;
      const int64_t ws = __shfl(s, w, 64), we = __shfl(e, w, 64);
      const int64_t wl = __shfl(lo, w, 64), wh = __shfl(hi, w, 64);
      const bool wi = __shfl((int)in_s, w, 64) != 0;
      const int wv = (int)(base + w);
      for (int64_t j = ws + lane; j < we; j += WAVE) {
        const int u = cols[j];
        if (!wi || u < wl || u >= wh) cc_link(wv, u, p);
      }
    }
  }
}

}  // namespace gsw
