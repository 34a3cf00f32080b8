// Device-resident direction-optimising BFS for one query.
//
// The engine's level loop (inspect -> select -> filter -> expand) reads the
// inspector's counters back to the host every level, and the host picks the
// next kernel. Here the whole choice stays on the device, so a query is a
// captured sequence of two kernels per level replayed as one hipGraph,
// with one host read per replay:
//   k_do_expand   one level: top-down (push) over the frontier bitmap, or
//                 bottom-up (pull) over the unvisited vertices, as the
//                 control block says; writes the next frontier bitmap and
//                 the next level's counters (vertices, out-edges)
//   k_do_decide   one workgroup: reduces the counters, stops the loop on an
//                 empty frontier and picks the next direction with the
//                 switching rule of Beamer, Asanovic and Patterson
//                 ("Direction-Optimizing Breadth-First Search", SC 2012):
//                 top-down -> bottom-up when the frontier's edges exceed
//                 the unexplored vertices' edges / alpha and the frontier
//                 grew, bottom-up -> top-down when the frontier holds fewer
//                 than n / beta vertices and shrank.
// Frontiers are bitmaps (one uint64 per 64 vertices) in three rotating
// buffers: level L reads F[L % 3], writes F[(L + 1) % 3] and clears
// F[(L + 2) % 3] for the level after. The level, the direction and the
// counters live in the control block, so every captured level is the same
// pair of launches.
//
// Top-down: wavefronts take 64 frontier words per step; for every non-zero
// word its vertices' edges are walked 64 at a time (degree scan across the
// lanes, owner found in LDS). A vertex of more than `heavy_deg` out-edges
// is not walked by one wavefront: every such vertex found in a level is
// appended to a short list, and the next top-down level splits the listed
// vertices' edges evenly over the whole grid (prefix of their degrees
// computed by k_do_decide). If the list overflows, its vertices are walked
// by their own wavefront, 64 edges per step.
// Bottom-up: one wavefront per frontier word, one lane per target; the
// first probe is the per-vertex copy of the first in-neighbour (coalesced),
// then the in-row until a frontier source is found (early exit). The next
// frontier word is the wavefront's ballot (plain store, every word written).
//
// The result is the BFS level array, the same whatever the directions.
#pragma once
#include "../utils/runtime.h"
#include "../utils/wave.h"

namespace gsw {

constexpr int DO_HCAP = 16384;   // listed heavy vertices per level
constexpr int DO_SLOTS = 64;     // partial-counter slots, one line each
constexpr int DO_STRIDE = 16;    // long longs per slot
constexpr int DO_DECIDE_BLOCK = 1024;
constexpr int DO_HIST = 4096;    // levels with a history row
constexpr int DO_HIST_W = 4;     // direction, frontier, edges, vertices
constexpr int DO_BU = 2;         // frontier words per wavefront step (bottom-up;
                                 // R-MAT 24 level 2: 1 word 178 us, 2 words 169 us,
                                 // 4 words 251 us: the arrays leave registers)

// control block (long long words)
enum : int {
  DO_LEVEL = 0,   // level being expanded (its frontier holds level L)
  DO_HALT,        // 1 once a level found nothing
  DO_DIR,         // 0 top-down, 1 bottom-up (for level DO_LEVEL)
  DO_NF,          // frontier vertices of level DO_LEVEL
  DO_MF,          // their out-edges
  DO_VISE,        // out-edges of every visited vertex
  DO_HN,          // listed heavy vertices of the frontier
  DO_HOVF,        // 1: the frontier's heavy list overflowed
  DO_HNEXT,       // appends to the next level's list
  DO_HE,          // out-edges of the listed heavy vertices
  DO_LEVELS,      // non-empty frontiers expanded (set on halt)
  DO_CTRL_WORDS = 16,
  DO_SLOT_BASE = DO_CTRL_WORDS,
  DO_TOTAL = DO_SLOT_BASE + DO_SLOTS * DO_STRIDE
};

struct dobfs_t {
  int64_t n = 0, nw = 0, m = 0;
  const int64_t *rows = nullptr;   // out-CSR
  const int32_t *cols = nullptr;
  const int64_t *rrows = nullptr;  // in-CSR
  const int32_t *rcols = nullptr;
  const int32_t *first = nullptr;  // first in-neighbour, -1 without in-edges
  const int32_t *odeg = nullptr;
  const int32_t *ideg = nullptr;
  int32_t *lv = nullptr;           // levels, -1 unreached
  uint64_t *vis = nullptr;         // [nw] visited
  uint64_t *F = nullptr;           // [3 * nw] rotating frontiers
  int32_t *heavy = nullptr;        // [2 * DO_HCAP] heavy lists (level parity)
  int64_t *hpre = nullptr;         // [DO_HCAP + 1] degree prefix of the list
  long long *ctrl = nullptr;       // [DO_TOTAL]
  long long *hist = nullptr;       // [DO_HIST][DO_HIST_W] per-level history
  int heavy_deg = 4096;
  float alpha = 14.f, beta = 24.f;
};

// Query reset: levels -1 (root 0), visited and frontier words zero except
// the root's bit, control block for level 0 (top-down from the root; the
// root is the heavy list of level 0 when its degree is above heavy_deg).
// root < 0: every vertex unreached and every word zero; the roots are then
// set by k_do_seed.
__global__ void __launch_bounds__(BLOCK) k_do_reset(dobfs_t d, int root) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t rw = root >> 6;
  const uint64_t rb = 1ull << (root & 63);
  for (int64_t v = tid; v < d.n; v += stride) d.lv[v] = v == root ? 0 : -1;
  for (int64_t w = tid; w < d.nw; w += stride) {
    const uint64_t x = w == rw ? rb : 0ull;
    d.vis[w] = x;
    d.F[w] = x;
    d.F[d.nw + w] = 0ull;
    d.F[2 * d.nw + w] = 0ull;
  }
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < DO_TOTAL; i += blockDim.x) d.ctrl[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0 && root >= 0) {
      const long long deg = d.odeg[root];
      d.ctrl[DO_NF] = 1;
      d.ctrl[DO_MF] = deg;
      d.ctrl[DO_VISE] = deg;
      if (deg > d.heavy_deg) {
        d.heavy[0] = root;
        d.hpre[0] = 0;
        d.hpre[1] = deg;
        d.ctrl[DO_HN] = 1;
        d.ctrl[DO_HE] = deg;
      }
    }
  }
}

// Several roots (distinct ids, after k_do_reset with root < 0): level 0,
// their bits in the visited and level-0 frontier words, and the level's
// counters. Roots of more than heavy_deg edges are not listed: the heavy
// list is marked overflowed, so each is walked by its own wavefront.
__global__ void __launch_bounds__(BLOCK)
    k_do_seed(dobfs_t d, const int32_t *__restrict__ roots, int64_t k) {
  __shared__ long long sh[BLOCK / WAVE];
  long long nf = 0, mf = 0;
  bool heavy = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int r = roots[i];
    d.lv[r] = 0;
    const unsigned long long b = 1ull << (r & 63);
    atomicOr((unsigned long long *)&d.vis[r >> 6], b);
    atomicOr((unsigned long long *)&d.F[r >> 6], b);
    const int deg = d.odeg[r];
    ++nf;
    mf += deg;
    heavy |= deg > d.heavy_deg;
  }
  const long long a = block_sum(nf, sh);
  const long long c = block_sum(mf, sh);
  const int hv = __syncthreads_or(heavy);
  if (threadIdx.x == 0) {
    if (a) atomicAdd((unsigned long long *)&d.ctrl[DO_NF], (unsigned long long)a);
    if (c) {
      atomicAdd((unsigned long long *)&d.ctrl[DO_MF], (unsigned long long)c);
      atomicAdd((unsigned long long *)&d.ctrl[DO_VISE], (unsigned long long)c);
    }
    if (hv) atomicOr((unsigned long long *)&d.ctrl[DO_HOVF], 1ull);
  }
}

struct do_level_t {
  int64_t L;
  const uint64_t *Fc;
  uint64_t *Fn;
  int32_t *hnext;
};

// A vertex found for level L + 1: level, counters, heavy list.
__device__ __forceinline__ void do_found(const dobfs_t &d, const do_level_t &t, int u,
                                         long long &nf, long long &mf) {
  d.lv[u] = (int)(t.L + 1);
  const int du = d.odeg[u];
  ++nf;
  mf += du;
  if (du > d.heavy_deg) {
    const long long i = atomicAdd((unsigned long long *)&d.ctrl[DO_HNEXT], 1ull);
    if (i < DO_HCAP) t.hnext[i] = u;
  }
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t x, int src) {
  const int lo = __shfl((int)(uint32_t)x, src, 64);
  const int hi = __shfl((int)(uint32_t)(x >> 32), src, 64);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Top-down visits of a whole wavefront (every lane calls it; `valid` lanes
// carry a target u). Lanes walk sorted adjacency, so consecutive lanes often
// hit the same visited word: each run of consecutive lanes with one word
// ORs its still-unvisited bits (segmented scan over the run) and its last
// lane sets them with one atomicOr on the visited word and one on the next
// frontier word, instead of two atomics per edge. A lane owns its target
// if the word's old value lacks the bit and no lower lane of the run has
// the same target (a target listed by two frontier vertices).
__device__ __forceinline__ void do_visit_wave(const dobfs_t &d, const do_level_t &t,
                                              bool valid, int u, long long &nf,
                                              long long &mf) {
  const int lane = lane_id();
  const int64_t w = valid ? (int64_t)(u >> 6) : -1;
  const uint64_t b = valid ? 1ull << (u & 63) : 0ull;
  const bool cand = valid && !(d.vis[w] & b);
  if (!ballot(cand)) return;
  // runs of equal words among consecutive lanes
  const long long wp = __shfl_up((long long)w, 1, 64);
  const bool head = lane == 0 || wp != w;
  const uint64_t heads = ballot(head);
  const int start = 63 - __clzll((long long)(heads & ((2ull << lane) - 1)));
  uint64_t x = cand ? b : 0ull;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint64_t y = __shfl_up(x, o, 64);
    if (lane - o >= start) x |= y;
  }
  // x: OR of the run's candidate bits up to this lane
  const uint64_t up = __shfl_up(x, 1, 64);
  const uint64_t below = lane > start ? up : 0ull;
  const bool tail = lane == WAVE - 1 || ((heads >> (lane + 1)) & 1ull);
  uint64_t old = 0;
  if (tail && x) {
    old = atomicOr((unsigned long long *)&d.vis[w], (unsigned long long)x);
    const uint64_t fresh = x & ~old;
    if (fresh) atomicOr((unsigned long long *)&t.Fn[w], (unsigned long long)fresh);
  }
  const uint64_t tails = ballot(tail);
  const int my_tail = __builtin_ctzll(tails & ~((1ull << lane) - 1));
  old = __shfl(old, my_tail, 64);
  if (cand && !(old & b) && !(below & b)) do_found(d, t, u, nf, mf);
}

__global__ void __launch_bounds__(BLOCK) k_do_expand(dobfs_t d) {
  __shared__ int s_scan[BLOCK];
  __shared__ int64_t s_start[BLOCK];
  __shared__ long long sh[BLOCK / WAVE];
  const long long *ctrl = d.ctrl;
  if (ctrl[DO_HALT]) return;
  do_level_t t;
  t.L = ctrl[DO_LEVEL];
  const bool pull = ctrl[DO_DIR] != 0;
  t.Fc = d.F + (t.L % 3) * d.nw;
  t.Fn = d.F + ((t.L + 1) % 3) * d.nw;
  uint64_t *Fz = d.F + ((t.L + 2) % 3) * d.nw;
  t.hnext = d.heavy + ((t.L + 1) & 1) * DO_HCAP;
  const int lane = lane_id();
  const int off = threadIdx.x & ~(WAVE - 1);
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t wave = tid >> LOG_WAVE;
  const int64_t nwaves = stride >> LOG_WAVE;
  long long nf = 0, mf = 0;  // next frontier: vertices, out-edges
  long long ne = 0, nv = 0;  // this level: edges inspected, vertices expanded
  for (int64_t i = tid; i < d.nw; i += stride) Fz[i] = 0ull;

  if (!pull) {
    const bool hovf = ctrl[DO_HOVF] != 0;
    // listed heavy vertices: their edges split evenly over the grid
    const int64_t HE = hovf ? 0 : ctrl[DO_HE];
    if (HE > 0) {
      const int hn = (int)ctrl[DO_HN];
      const int32_t *hcur = d.heavy + (t.L & 1) * DO_HCAP;
      for (int64_t k0 = tid - lane; k0 < HE; k0 += stride) {
        const int64_t k = k0 + lane;
        const bool ok = k < HE;
        int u = 0;
        if (ok) {
          ++ne;
          const int o = owner_search(d.hpre, hn, k);
          const int v = hcur[o];
          u = d.cols[d.rows[v] + (k - d.hpre[o])];
        }
        do_visit_wave(d, t, ok, u, nf, mf);
      }
    }
    // the other frontier vertices: 64 words per wavefront step
    for (int64_t w0 = wave * WAVE; w0 < d.nw; w0 += nwaves * WAVE) {
      const int64_t wi = w0 + lane;
      const uint64_t word = wi < d.nw ? t.Fc[wi] : 0ull;
      uint64_t nz = ballot(word != 0ull);
      while (nz) {
        const int j = __builtin_ctzll(nz);
        nz &= nz - 1;
        const uint64_t bits = shfl_u64(word, j);
        const int v = (int)(((w0 + j) << 6) + lane);
        int deg = 0;
        int64_t start = 0;
        bool alone = false;
        if ((bits >> lane) & 1ull) {
          ++nv;
          deg = d.odeg[v];
          if (deg > d.heavy_deg) {
            alone = hovf;  // walked below by this wavefront
            deg = 0;       // otherwise walked by the heavy pass
          } else {
            start = d.rows[v];
          }
        }
        ne += deg;
        int tot;
        const int ex = wave_exclusive_scan(deg, tot);
        s_scan[off + lane] = ex;
        s_start[off + lane] = start;
        wave_sync();
        for (int j0 = 0; j0 < tot; j0 += WAVE) {
          const int k = j0 + lane;
          int u = 0;
          if (k < tot) {
            const int o = owner_search(&s_scan[off], WAVE, k);
            u = d.cols[s_start[off + o] + (k - s_scan[off + o])];
          }
          do_visit_wave(d, t, k < tot, u, nf, mf);
        }
        wave_sync();
        // heavy vertices of an overflowed list: one wavefront each
        uint64_t al = ballot(alone);
        while (al) {
          const int a = __builtin_ctzll(al);
          al &= al - 1;
          const int va = __shfl(v, a, 64);
          const int64_t s = d.rows[va], e = d.rows[va + 1];
          for (int64_t k0 = s; k0 < e; k0 += WAVE) {
            const int64_t k = k0 + lane;
            const bool ok = k < e;
            if (ok) ++ne;
            do_visit_wave(d, t, ok, ok ? d.cols[k] : 0, nf, mf);
          }
        }
      }
    }
  } else {
    // bottom-up: DO_BU consecutive frontier words per wavefront step, their
    // first probes issued together (a probe is a chain of dependent loads)
    for (int64_t w0 = wave * DO_BU; w0 < d.nw; w0 += nwaves * DO_BU) {
      uint64_t vw[DO_BU];
      int id[DO_BU], fu[DO_BU];
#pragma unroll
      for (int j = 0; j < DO_BU; ++j) vw[j] = w0 + j < d.nw ? d.vis[w0 + j] : ~0ull;
#pragma unroll
      for (int j = 0; j < DO_BU; ++j) {
        const int64_t v = ((w0 + j) << 6) + lane;
        const bool act = v < d.n && !((vw[j] >> lane) & 1ull);
        id[j] = act ? d.ideg[v] : 0;
        fu[j] = act && id[j] > 0 ? d.first[v] : -1;
        if (act) ++nv;
      }
#pragma unroll
      for (int j = 0; j < DO_BU; ++j) {
        const int64_t w = w0 + j;
        if (w >= d.nw) break;  // uniform
        const int64_t v = (w << 6) + lane;
        bool found = false;
        int u = fu[j];
        if (u >= 0) {
          ++ne;
          found = (t.Fc[u >> 6] >> (u & 63)) & 1ull;
          if (!found && id[j] > 1) {
            const int64_t s = d.rrows[v];
            for (int k = 1; k < id[j]; ++k) {
              u = d.rcols[s + k];
              ++ne;
              if ((t.Fc[u >> 6] >> (u & 63)) & 1ull) {
                found = true;
                break;
              }
            }
          }
        }
        const uint64_t fm = ballot(found);
        if (lane == 0) {
          t.Fn[w] = fm;
          if (fm) d.vis[w] = vw[j] | fm;
        }
        if (found) do_found(d, t, (int)v, nf, mf);
      }
    }
  }
  // counters of level L + 1 into this workgroup's slot
  long long *slot = d.ctrl + DO_SLOT_BASE + (int64_t)(blockIdx.x & (DO_SLOTS - 1)) * DO_STRIDE;
  const long long a = block_sum(nf, sh);
  if (threadIdx.x == 0 && a) atomicAdd((unsigned long long *)&slot[0], (unsigned long long)a);
  const long long b = block_sum(mf, sh);
  if (threadIdx.x == 0 && b) atomicAdd((unsigned long long *)&slot[1], (unsigned long long)b);
  const long long c = block_sum(ne, sh);
  if (threadIdx.x == 0 && c) atomicAdd((unsigned long long *)&slot[2], (unsigned long long)c);
  const long long e = block_sum(nv, sh);
  if (threadIdx.x == 0 && e) atomicAdd((unsigned long long *)&slot[3], (unsigned long long)e);
}

// End of level L: counters of level L + 1, halt on an empty frontier, the
// direction of level L + 1, its heavy list and the list's degree prefix.
__global__ void __launch_bounds__(DO_DECIDE_BLOCK) k_do_decide(dobfs_t d) {
  __shared__ long long s_red[4][DO_DECIDE_BLOCK / WAVE];
  __shared__ int s_go;
  __shared__ long long s_w[DO_DECIDE_BLOCK / WAVE];
  long long *ctrl = d.ctrl;
  if (ctrl[DO_HALT]) return;
  long long nf = 0, mf = 0, ne = 0, nv = 0;
  if (threadIdx.x < DO_SLOTS) {
    long long *slot = ctrl + DO_SLOT_BASE + threadIdx.x * DO_STRIDE;
    nf = slot[0];
    mf = slot[1];
    ne = slot[2];
    nv = slot[3];
    slot[0] = slot[1] = slot[2] = slot[3] = 0;
  }
  nf = block_sum(nf, s_red[0]);
  mf = block_sum(mf, s_red[1]);
  ne = block_sum(ne, s_red[2]);
  nv = block_sum(nv, s_red[3]);
  const int nth = DO_DECIDE_BLOCK;
  if (threadIdx.x == 0) {
    const long long L = ctrl[DO_LEVEL];
    s_go = 0;
    if (L < DO_HIST) {
      long long *hrow = d.hist + L * DO_HIST_W;
      hrow[0] = ctrl[DO_DIR];
      hrow[1] = ctrl[DO_NF];
      hrow[2] = ne;
      hrow[3] = nv;
    }
    if (nf == 0) {
      ctrl[DO_HALT] = 1;
      ctrl[DO_LEVELS] = L + 1;
    } else {
      const long long vise = ctrl[DO_VISE] + mf;
      const long long mu = d.m - vise;
      const long long prev_nf = ctrl[DO_NF];
      const bool was_pull = ctrl[DO_DIR] != 0;
      bool pull;
      if (!was_pull)
        pull = (double)mf > (double)mu / d.alpha && nf > prev_nf;
      else
        pull = !((double)nf < (double)d.n / d.beta && nf < prev_nf);
      const long long hn = ctrl[DO_HNEXT];
      ctrl[DO_VISE] = vise;
      ctrl[DO_NF] = nf;
      ctrl[DO_MF] = mf;
      ctrl[DO_DIR] = pull ? 1 : 0;
      ctrl[DO_HN] = hn < DO_HCAP ? hn : DO_HCAP;
      ctrl[DO_HOVF] = hn > DO_HCAP ? 1 : 0;
      ctrl[DO_HNEXT] = 0;
      ctrl[DO_HE] = 0;
      ctrl[DO_LEVEL] = L + 1;
      s_go = (!pull && hn > 0 && hn <= DO_HCAP) ? 1 : 0;
    }
  }
  __syncthreads();
  if (!s_go) return;
  // degree prefix of level L + 1's heavy list (the list written in level L)
  const long long L1 = ctrl[DO_LEVEL];
  const int hn = (int)ctrl[DO_HN];
  const int32_t *h = d.heavy + (L1 & 1) * DO_HCAP;
  constexpr int PER = DO_HCAP / DO_DECIDE_BLOCK;
  long long loc[PER];
  long long sum = 0;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int k = threadIdx.x * PER + i;
    loc[i] = k < hn ? (long long)d.odeg[h[k]] : 0;
    sum += loc[i];
  }
  // exclusive scan of the per-thread sums in 64 bits: shuffles inside the
  // wavefront, wavefront totals through LDS
  long long incl = sum;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const long long y = __shfl_up(incl, o, 64);
    if (lane_id() >= o) incl += y;
  }
  long long ex = incl - sum;
  const int w = threadIdx.x >> LOG_WAVE;
  if (lane_id() == WAVE - 1) s_w[w] = incl;
  __syncthreads();
  long long tot = 0;
  for (int i = 0; i < nth / WAVE; ++i) {
    ex += i < w ? s_w[i] : 0;
    tot += s_w[i];
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int k = threadIdx.x * PER + i;
    if (k < hn) d.hpre[k] = ex;
    ex += loc[i];
  }
  if (threadIdx.x == 0) {
    d.hpre[hn] = tot;
    ctrl[DO_HE] = tot;
  }
}

}  // namespace gsw
