// Push (top-down) expand kernels.
//
// The edge body is the reference's (standalone: __expand_VC_{TM,WM,CM,
// STRICT}; fused: *_fused / *_wtf in src/kernel_libs/expand_VC_*.cuh):
//   standalone: [pruning: first visit of u] -> [not to-all: u inactive]
//               -> compAtomic(wa(u), emit(v, e))
//   fused:      [not to-all: cond(u, msg)] -> compAtomic -> exact de-dup on
//               the visited bitmap -> append u to the next queue
// What differs per load-balancing strategy is only how (v, u, e) triples are
// spread over lanes. Strategies, rebuilt for 64-lane wavefronts:
//   TM     thread per frontier vertex, serial edge loop
//   WM     wavefront per 64 frontier vertices: DPP scan of degrees, LDS-staged
//          (vertex, offset, prefix), 64 edges per step, owner found by binary
//          search over the 64 prefixes in LDS
//   CM     workgroup per 256 frontier vertices, same scheme with a block scan
//   STRICT edge-balanced: int64 prefix of (deg + 1) over the frontier, every
//          workgroup owns an equal slice of slots; per 256-slot chunk the 257
//          covering prefixes are staged in LDS (the +1 guarantees a chunk
//          never spans more than 257 vertices); the owner of the next chunk
//          start is found in the current window, so only a workgroup's first
//          chunk searches global memory (merge-path style)
//   TWC    degree bins: small -> thread, medium -> wavefront,
//          large -> workgroup, all three processed in one launch
#pragma once
#include "../abstraction/config.h"
#include "../data_structures/active_set.h"
#include "../data_structures/functor.h"
#include "../utils/wave.h"
#include "scan.h"

namespace gsw {

// Edge / vertex work counters of an expand: a workgroup reduction, then one
// atomic per counter into the workgroup's slot of the current record row
// (slot blockIdx.x % WORK_SLOTS; the slots of a row lie WORK_ROWS rows
// apart, so each has a cache line of its own: atomics from every workgroup
// onto one line serialise on MI355X). Every thread of the workgroup calls
// it; a workgroup that leaves early leaves as a whole and counts nothing.
template <int NT = BLOCK>
__device__ __forceinline__ void add_work(unsigned long long *work,
                                         unsigned long long edges,
                                         unsigned long long verts) {
  __shared__ unsigned long long sh[2][NT / WAVE];
  edges = wave_sum_any(edges);
  verts = wave_sum_any(verts);
  if (lane_id() == 0) {
    sh[0][threadIdx.x >> LOG_WAVE] = edges;
    sh[1][threadIdx.x >> LOG_WAVE] = verts;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long e = 0, v = 0;
    const int nw = (blockDim.x + WAVE - 1) >> LOG_WAVE;
    for (int i = 0; i < nw; ++i) {
      e += sh[0][i];
      v += sh[1][i];
    }
    unsigned long long *slot =
        work + (int64_t)(blockIdx.x & (WORK_SLOTS - 1)) * WORK_SLOT_STRIDE;
    if (e) atomicAdd(slot + WORK_EDGES, e);
    if (v) atomicAdd(slot + WORK_VERTS, v);
  }
}

template <typename G, typename F, bool FUSED> struct push_body {
  static constexpr bool fused = FUSED;
  active_set_t as;
  G g;
  F f;
  config_t conf;

  __device__ __forceinline__ const int64_t *rows() const { return g.dg_row_ptr; }
  __device__ __forceinline__ const int32_t *cols() const { return g.dg_adj_list; }

  // true when u must be appended to the next queue (fused only)
  __device__ __forceinline__ bool operator()(int v, int u, int64_t e) {
    if constexpr (!FUSED) {
      bool go = true;
      if (conf.conf_pruning) go = as.bitmap.mark_duplicate_lite(u);
      if (go && !conf.conf_toall) go = as.bitmap.is_inactive(u);
      if (go) {
        auto msg = f.emit(v, g.fetch_edata(e), g);
        f.compAtomic(f.wa_of(u), msg, g);
      }
      return false;
    } else {
      if (conf.conf_pruning && as.bitmap.visited.test(u)) return false;
      auto msg = f.emit(v, g.fetch_edata(e), g);
      bool go = true;
      if (!conf.conf_toall) go = f.cond(u, msg, g);
      if (go) go = f.compAtomic(f.wa_of(u), msg, g);
      if (go) go = as.bitmap.mark_duplicate_atomic(u);
      return go;
    }
  }

  // fused pushes stage the next queue per wavefront in LDS (STAGE_LDS ints
  // per workgroup: the wavefronts' buffers, then the workgroup's flush
  // counts); emit_q is called by all 64 lanes of a wavefront, emit_flush by
  // every thread of the workgroup at the end of the kernel
  using stage_t = queue_t::stage_t;
  static constexpr int STAGE_WAVES = BLOCK / WAVE;
  static constexpr int STAGE_LDS = FUSED ? STAGE_WAVES * queue_t::STAGE + STAGE_WAVES + 1 : 1;
  __device__ __forceinline__ stage_t stage(int *lds) const {
    stage_t st;
    st.buf = FUSED ? lds + (threadIdx.x >> LOG_WAVE) * queue_t::STAGE : lds;
    return st;
  }
  __device__ __forceinline__ void emit_q(bool q, int u, stage_t &st) {
    if constexpr (FUSED) as.queue.push(q, u, st);
  }
  __device__ __forceinline__ void emit_flush(stage_t &st, int *lds) {
    if constexpr (FUSED)
      as.queue.flush_stage_block(st, lds + STAGE_WAVES * queue_t::STAGE);
  }
};

// ---------------------------------------------------------------- TM
template <class B>
__global__ void __launch_bounds__(BLOCK) k_push_tm(B body) {
  __shared__ int s_stage[B::STAGE_LDS];
  auto st = body.stage(s_stage);
  const int lane = lane_id();
  const int64_t N = body.as.get_size();
  const Status want = body.conf.want();
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  unsigned long long ne = 0, nv = 0;
  // the loops are uniform over the wavefront (staged emits need all lanes)
  for (int64_t base = tid - lane; base < N; base += stride) {
    const int64_t idx = base + lane;
    const int v = (idx < N) ? body.as.fetch(idx, want) : -1;
    int64_t j = 0, e = 0;
    if (v >= 0) {
      j = body.rows()[v];
      e = body.rows()[v + 1];
      ne += e - j;
      ++nv;
    }
    while (ballot(j < e)) {
      int u = -1;
      bool q = false;
      if (j < e) {
        u = body.cols()[j];
        q = body(v, u, j);
        ++j;
      }
      body.emit_q(q, u, st);
    }
  }
  body.emit_flush(st, s_stage);
  add_work(body.as.dg_work, ne, nv);
}

// ---------------------------------------------------------------- WM
template <class B>
// A wavefront takes `vpw` frontier items (64, 32, 16 or 8) and walks their
// edges 64 at a time: with a small mean degree fewer items per wavefront
// means more wavefronts in flight and fewer dependent 64-edge steps each.
__global__ void __launch_bounds__(BLOCK) k_push_wm(B body, int vpw) {
  __shared__ int s_stage[B::STAGE_LDS];
  auto st = body.stage(s_stage);
  __shared__ int s_v[BLOCK];
  __shared__ int s_scan[BLOCK];
  __shared__ int64_t s_start[BLOCK];
  const int lane = lane_id();
  const int off = threadIdx.x & ~(WAVE - 1);
  const int64_t N = body.as.get_size();
  const Status want = body.conf.want();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  unsigned long long ne = 0, nv = 0;
  for (int64_t base = wave * vpw; base < N; base += nwaves * vpw) {
    const int64_t idx = base + lane;
    const int v = (lane < vpw && idx < N) ? body.as.fetch(idx, want) : -1;
    int64_t start = 0;
    int deg = 0;
    if (v >= 0) {
      start = body.rows()[v];
      deg = (int)(body.rows()[v + 1] - start);
      ++nv;
    }
    int tot;
    const int ex = wave_exclusive_scan(deg, tot);
    s_v[off + lane] = v;
    s_scan[off + lane] = ex;
    s_start[off + lane] = start;
    wave_sync();
    ne += deg;
    for (int j0 = 0; j0 < tot; j0 += WAVE) {
      const int j = j0 + lane;
      int u = -1;
      bool q = false;
      if (j < tot) {
        const int o = owner_search(&s_scan[off], WAVE, j);
        const int64_t e = s_start[off + o] + (j - s_scan[off + o]);
        u = body.cols()[e];
        q = body(s_v[off + o], u, e);
      }
      body.emit_q(q, u, st);
    }
    wave_sync();
  }
  body.emit_flush(st, s_stage);
  add_work(body.as.dg_work, ne, nv);
}

// ---------------------------------------------------------------- CM
template <class B>
__global__ void __launch_bounds__(BLOCK) k_push_cm(B body) {
  __shared__ int s_stage[B::STAGE_LDS];
  auto st = body.stage(s_stage);
  __shared__ int s_v[BLOCK];
  __shared__ int s_scan[BLOCK];
  __shared__ int64_t s_start[BLOCK];
  __shared__ int s_red[BLOCK / WAVE];
  const int64_t N = body.as.get_size();
  const Status want = body.conf.want();
  unsigned long long ne = 0, nv = 0;
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < N;
       base += (int64_t)gridDim.x * BLOCK) {
    const int64_t idx = base + threadIdx.x;
    const int v = (idx < N) ? body.as.fetch(idx, want) : -1;
    int64_t start = 0;
    int deg = 0;
    if (v >= 0) {
      start = body.rows()[v];
      deg = (int)(body.rows()[v + 1] - start);
      ++nv;
    }
    int tot;
    const int ex = block_exclusive_scan(deg, tot, s_red);
    s_v[threadIdx.x] = v;
    s_scan[threadIdx.x] = ex;
    s_start[threadIdx.x] = start;
    __syncthreads();
    ne += deg;
    for (int j0 = 0; j0 < tot; j0 += BLOCK) {
      const int j = j0 + threadIdx.x;
      int u = -1;
      bool q = false;
      if (j < tot) {
        const int o = owner_search(s_scan, BLOCK, j);
        const int64_t e = s_start[o] + (j - s_scan[o]);
        u = body.cols()[e];
        q = body(s_v[o], u, e);
      }
      body.emit_q(q, u, st);
    }
    __syncthreads();
  }
  body.emit_flush(st, s_stage);
  add_work(body.as.dg_work, ne, nv);
}

// ---------------------------------------------------------------- STRICT
// slot counts (deg + 1) of the frontier items; pull selects in-degrees
template <typename G>
__global__ void __launch_bounds__(BLOCK)
    k_strict_prepare(active_set_t as, G g, Status want, bool pull,
                     int64_t *d) {
  const int64_t N = as.get_size();
  const int64_t *rows = pull ? g.dgr_row_ptr : g.dg_row_ptr;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int v = as.fetch(i, want);
    d[i] = 1 + ((v >= 0) ? (rows[v + 1] - rows[v]) : 0);
  }
}

// Slot prefix of the STRICT partition: an explicit int64 scan of (deg + 1)
// over the frontier items (queue frontiers), or, for bitmap frontiers where
// item k is vertex k, the implicit prefix rows[k] + k (no scan launches).
struct slots_t {
  const int64_t *scan;  // explicit prefix, or nullptr
  const int64_t *rows;  // offsets of the traversed direction
  __device__ __forceinline__ int64_t at(int64_t k) const {
    return scan ? scan[k] : rows[k] + k;
  }
  // largest k in [0, len) with at(k) <= x (at(0) <= x)
  __device__ __forceinline__ int64_t last_le(int64_t len, int64_t x) const {
    int64_t lo = 0, hi = len;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (at(mid) <= x)
        lo = mid;
      else
        hi = mid;
    }
    return lo;
  }
};

// A STRICT chunk is SITEMS slots per lane: slots c + j*BLOCK + lane, so each
// lane keeps SITEMS independent loads in flight and every load instruction
// of a wavefront stays coalesced. The window holds the SCHUNK + 1 prefixes
// that can own the chunk's slots (each item owns >= 1 slot).
constexpr int SITEMS = 4;
constexpr int SCHUNK = BLOCK * SITEMS;

// Stage the SCHUNK + 1 prefixes starting at item k0 (clamped to N) into LDS.
template <class B>
__device__ __forceinline__ void strict_window(B &body, const slots_t &sl,
                                              int64_t N, int64_t k0,
                                              Status want, int64_t *s_scan,
                                              int *s_v, int64_t *s_start,
                                              const int32_t *tgt = nullptr) {
  // tgt (compressed rows): item k is row k of `rows`, target vertex tgt[k]
  for (int t = threadIdx.x; t <= SCHUNK; t += blockDim.x) {
    const int64_t k = k0 + t;
    s_scan[t] = sl.at(k < N ? k : N);
    int v = -1;
    int64_t st = 0;
    if (k < N) {
      v = body.as.fetch(tgt ? (int64_t)tgt[k] : k, want);
      if (v >= 0) st = sl.rows[tgt ? k : v];
    }
    s_v[t] = v;
    s_start[t] = st;
  }
}

// Owner of the next chunk start: inside the staged window, or exactly the
// item right after it (every item owns at least one slot).
__device__ __forceinline__ int64_t strict_next_owner(const int64_t *s_scan,
                                                     int64_t k0, int64_t nc) {
  return (nc < s_scan[SCHUNK]) ? k0 + owner_search(s_scan, SCHUNK + 1, nc)
                               : k0 + SCHUNK;
}

template <class B>
__global__ void __launch_bounds__(BLOCK)
    k_push_strict(B body, const int64_t *__restrict__ scan) {
  __shared__ int s_stage[B::STAGE_LDS];
  auto st = body.stage(s_stage);
  __shared__ int64_t s_scan[SCHUNK + 1];
  __shared__ int64_t s_start[SCHUNK + 1];
  __shared__ int s_v[SCHUNK + 1];
  const slots_t sl{scan, body.rows()};
  const int64_t N = body.as.get_size();
  const Status want = body.conf.want();
  const int64_t total = sl.at(N);
  const int64_t per = (total + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = MIN(total, lo + per);
  unsigned long long ne = 0;
  if (lo >= hi) return;
  int64_t k0 = sl.last_le(N + 1, lo);
  for (int64_t c = lo; c < hi; c += SCHUNK) {
    strict_window(body, sl, N, k0, want, s_scan, s_v, s_start);
    __syncthreads();
    int u[SITEMS], v[SITEMS];
    int64_t e[SITEMS];
#pragma unroll
    for (int j = 0; j < SITEMS; ++j) {
      const int64_t slot = c + j * BLOCK + threadIdx.x;
      v[j] = -1;
      if (slot < hi) {
        const int o = owner_search(s_scan, SCHUNK + 1, slot);
        const int64_t r = slot - s_scan[o];
        const int64_t deg = s_scan[o + 1] - s_scan[o] - 1;
        if (s_v[o] >= 0 && r < deg) {
          v[j] = s_v[o];
          e[j] = s_start[o] + r;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < SITEMS; ++j) u[j] = (v[j] >= 0) ? body.cols()[e[j]] : -1;
#pragma unroll
    for (int j = 0; j < SITEMS; ++j) {
      bool q = false;
      if (v[j] >= 0) {
        q = body(v[j], u[j], e[j]);
        ++ne;
      }
      body.emit_q(q, u[j], st);
    }
    const int64_t nk = strict_next_owner(s_scan, k0, c + SCHUNK);
    __syncthreads();
    k0 = nk;
  }
  body.emit_flush(st, s_stage);
  add_work(body.as.dg_work, ne, 0);
}

// ---------------------------------------------------------------- TWC
template <typename G>
__global__ void __launch_bounds__(BLOCK)
    k_twc_bin(active_set_t as, G g, Status want, bool pull, int small,
              int large) {
  // one LDS stage per bin and wavefront: one atomic per bin per flush
  constexpr int CAP = 256;
  __shared__ int s_stage[3][BLOCK / WAVE][CAP];
  wave_stage_t<CAP> st[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) st[k].buf = s_stage[k][threadIdx.x >> LOG_WAVE];
  const int64_t N = as.get_size();
  const int64_t *rows = pull ? g.dgr_row_ptr : g.dg_row_ptr;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t cap = as.bins.cap;  // bins hold every frontier item
  for (int64_t i0 = tid - lane_id(); i0 < N; i0 += stride) {
    const int64_t i = i0 + lane_id();
    int v = (i < N) ? as.fetch(i, want) : -1;
    int b = -1;
    if (v >= 0) {
      const int64_t d = rows[v + 1] - rows[v];
      b = (d < small) ? 0 : (d < large) ? 1 : 2;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
      st[k].put(b == k, v, as.bins.sizes + k, as.bins.bin[k], cap,
                as.queue.overflow);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    st[k].flush(as.bins.sizes + k, as.bins.bin[k], cap, as.queue.overflow);
}

template <class B>
__global__ void __launch_bounds__(BLOCK) k_push_twc(B body) {
  __shared__ int s_stage[B::STAGE_LDS];
  auto st = body.stage(s_stage);
  const int lane = lane_id();
  const int32_t *sizes = body.as.bins.sizes;
  unsigned long long ne = 0, nv = 0;
  // large: one workgroup per vertex
  const int nl = sizes[2];
  for (int i = blockIdx.x; i < nl; i += gridDim.x) {
    const int v = body.as.bins.bin[2][i];
    const int64_t s = body.rows()[v], e = body.rows()[v + 1];
    if (threadIdx.x == 0) {
      ne += e - s;
      ++nv;
    }
    for (int64_t j0 = s; j0 < e; j0 += BLOCK) {
      const int64_t j = j0 + threadIdx.x;
      int u = -1;
      bool q = false;
      if (j < e) {
        u = body.cols()[j];
        q = body(v, u, j);
      }
      body.emit_q(q, u, st);
    }
  }
  // medium: one wavefront per vertex
  const int nm = sizes[1];
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  for (int64_t i = wave; i < nm; i += nwaves) {
    const int v = body.as.bins.bin[1][i];
    const int64_t s = body.rows()[v], e = body.rows()[v + 1];
    if (lane == 0) {
      ne += e - s;
      ++nv;
    }
    for (int64_t j0 = s; j0 < e; j0 += WAVE) {
      const int64_t j = j0 + lane;
      int u = -1;
      bool q = false;
      if (j < e) {
        u = body.cols()[j];
        q = body(v, u, j);
      }
      body.emit_q(q, u, st);
    }
  }
  // small: one thread per vertex (loops uniform over the wavefront)
  const int ns = sizes[0];
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = tid - lane; base < ns; base += stride) {
    const int64_t i = base + lane;
    int v = -1;
    int64_t j = 0, e = 0;
    if (i < ns) {
      v = body.as.bins.bin[0][i];
      j = body.rows()[v];
      e = body.rows()[v + 1];
      ne += e - j;
      ++nv;
    }
    while (ballot(j < e)) {
      int u = -1;
      bool q = false;
      if (j < e) {
        u = body.cols()[j];
        q = body(v, u, j);
        ++j;
      }
      body.emit_q(q, u, st);
    }
  }
  body.emit_flush(st, s_stage);
  add_work(body.as.dg_work, ne, nv);
}

}  // namespace gsw
