// Pull (bottom-up) expand kernels.
//
// Edge body as in the reference's __rexpand_VC_* kernels
// (src/kernel_libs/expand_VC_{TM,WM,CM,STRICT}.cuh): for a target v and an
// in-edge u -> v, if the source is active (or the algorithm is from-all) the
// message emit(u, e) is combined with comp; idempotent algorithms (pruning)
// stop at the first contributing in-edge (early exit).
//   TM     thread per target, serial in-edge loop, combined in registers
//   WM     wavefront per 64 targets, 64 in-edges per step; messages of one
//          target are merged by a segmented shuffle reduction and kept in LDS
//          until the wavefront moves on, then applied with comp
//   CM     workgroup per 256 targets; per-wavefront segments are applied with
//          compAtomic (a target may span wavefronts)
//   STRICT edge-balanced over the targets' in-edges (same partition as the
//          push STRICT); segments applied with compAtomic; early exit through
//          the visited bitmap
#pragma once
#include "expand_push.h"
#include "inspect.h"

namespace gsw {

template <typename T> __device__ __forceinline__ T shfl_down_any(T x, int d) {
  constexpr int W = (int)((sizeof(T) + 3) / 4);
  union U {
    T t;
    int i[W];
    __device__ U() {}
  } a, b;
  a.t = x;
#pragma unroll
  for (int k = 0; k < W; ++k) b.i[k] = __shfl_down(a.i[k], d, 64);
  return b.t;
}

template <typename T> __device__ __forceinline__ T shfl_xor_any(T x, int d) {
  constexpr int W = (int)((sizeof(T) + 3) / 4);
  union U {
    T t;
    int i[W];
    __device__ U() {}
  } a, b;
  a.t = x;
#pragma unroll
  for (int k = 0; k < W; ++k) b.i[k] = __shfl_xor(a.i[k], d, 64);
  return b.t;
}

template <typename G, typename F> struct pull_body {
  using WA = typename F::wa_t;
  active_set_t as;
  G g;
  F f;
  config_t conf;

  __device__ __forceinline__ const int64_t *rows() const { return g.dgr_row_ptr; }
  __device__ __forceinline__ const int32_t *cols() const { return g.dgr_adj_list; }
  // cols()[rows()[v]] for a target with in-edges, read from the per-vertex
  // copy (coalesced across a wavefront of consecutive targets)
  __device__ __forceinline__ int first_col(int v) const { return g.dgr_first[v]; }
  __device__ __forceinline__ bool source_ok(int u) const {
    return conf.conf_fromall || as.bitmap.is_active(u);
  }
  __device__ __forceinline__ WA msg(int u, int64_t e) {
    return f.emit(u, g.fetch_edata_r(e), g);
  }
  // Segmented reduction over lanes whose owner ids are non-decreasing:
  // afterwards the first lane of each owner run holds the combined message.
  __device__ __forceinline__ void seg_reduce(int owner, bool &vote, WA &m) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
      const WA om = shfl_down_any(m, d);
      const int oo = __shfl_down(owner, d, 64);
      const int ov = __shfl_down((int)vote, d, 64);
      if (lane + d < WAVE && oo == owner && ov) {
        if (vote)
          f.comp(&m, om, g);
        else
          m = om;
        vote = true;
      }
    }
  }
  __device__ __forceinline__ bool seg_head(int owner) const {
    const int prev = __shfl_up(owner, 1, 64);
    return owner >= 0 && (lane_id() == 0 || prev != owner);
  }
};

// ---------------------------------------------------------------- TM
template <class B>
__global__ void __launch_bounds__(BLOCK) k_pull_tm(B body) {
  using WA = typename B::WA;
  const int64_t N = body.as.get_size();
  const Status want = body.conf.want();
  const bool prune = body.conf.conf_pruning;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  unsigned long long ne = 0, nv = 0;
  for (int64_t idx = tid; idx < N; idx += stride) {
    const int v = body.as.fetch(idx, want);
    if (v < 0) continue;
    ++nv;
    const int64_t s = body.rows()[v], e = body.rows()[v + 1];
    WA acc;
    bool any = false;
    int64_t j = s, end = e;
    if (body.conf.pull_first && j < e) {
      // the first in-edge from the per-vertex copy; an early exit there
      // leaves the walk below empty
      const int u = body.first_col(v);
      if (body.source_ok(u)) {
        const WA m = body.msg(u, j);
        acc = *body.f.wa_of(v);
        any = true;
        body.f.comp(&acc, m, body.g);
        if (prune) end = j + 1;
      }
      ++j;
    }
    for (; j < end; ++j) {
      const int u = body.cols()[j];
      if (!body.source_ok(u)) continue;
      const WA m = body.msg(u, j);
      if (!any) {
        acc = *body.f.wa_of(v);
        any = true;
      }
      body.f.comp(&acc, m, body.g);
      if (prune) {
        ++j;
        break;
      }
    }
    ne += j - s;
    if (any) *body.f.wa_of(v) = acc;
  }
  add_work(body.as.dg_work, ne, nv);
}

// TM pull over a bitmap frontier with the next level's inspection folded
// in, for level-synchronous functors (a vertex updated in this level is
// exactly the next level's frontier; the others keep their status): one
// 64-vertex word per wavefront step, the ballot of updated vertices is the
// next `active` word (written to next_active, swapped in by the host), they
// leave the inactive set in place (each word is touched by its own
// wavefront only), and the inspector's six counters of the next level are
// reduced here, so that level needs no inspector pass.
template <class B>
__global__ void __launch_bounds__(BLOCK)
    k_pull_tm_inspect(B body, uint64_t *next_active, long long *slots) {
  using WA = typename B::WA;
  __shared__ long long sh[BLOCK / WAVE];
  const int lane = lane_id();
  const int64_t n = body.g.nvertexs;
  const int64_t nw = (n + 63) >> 6;
  const bool prune = body.conf.conf_pruning;
  uint64_t *inact_bits = body.as.bitmap.inactive.dg_bits;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  unsigned long long ne = 0, nv = 0;
  long long act = 0, inact = 0, pushw = 0, pullw = 0, maxa = 0, maxi = 0;
  for (int64_t w = wave; w < nw; w += nwaves) {
    const int64_t v = (w << 6) + lane;
    const uint64_t word = inact_bits[w];  // bit set: not inactive
    bool found = false;
    if (v < n && !((word >> lane) & 1ull)) {
      ++nv;
      // the row's offset is read only when an edge past the first copy is
      // walked (or the functor reads edge data): a target whose first
      // in-neighbour is in the frontier costs its degree and that copy
      const int64_t id = body.g.get_in_degree((int)v);
      int64_t s = -1;
      auto row = [&]() {
        if (s < 0) s = body.rows()[v];
        return s;
      };
      constexpr bool edata = decltype(body.g)::has_edata;
      WA acc;
      int64_t k = 0, end = id;  // positions within the row
      if (body.conf.pull_first && k < end) {
        const int u = body.first_col((int)v);
        if (body.source_ok(u)) {
          const WA m = body.msg(u, edata ? row() : 0);
          acc = *body.f.wa_of(v);
          found = true;
          body.f.comp(&acc, m, body.g);
          if (prune) end = k + 1;
        }
        ++k;
      }
      for (; k < end; ++k) {
        const int64_t j = row() + k;
        const int u = body.cols()[j];
        if (!body.source_ok(u)) continue;
        const WA m = body.msg(u, j);
        if (!found) {
          acc = *body.f.wa_of(v);
          found = true;
        }
        body.f.comp(&acc, m, body.g);
        if (prune) {
          ++k;
          break;
        }
      }
      ne += k;
      if (found) {
        *body.f.wa_of(v) = acc;
        const long long od = body.g.get_out_degree((int)v);
        ++act;
        pushw += od;
        maxa = od > maxa ? od : maxa;
      } else {
        ++inact;
        pullw += id;
        maxi = id > maxi ? id : maxi;
      }
    }
    const uint64_t fm = ballot(found);
    if (lane == 0) {
      next_active[w] = fm;
      if (fm) inact_bits[w] = word | fm;
    }
  }
  long long *slot = stat_slot(slots);
  flush_sum(act, sh, slot + 0);
  flush_sum(inact, sh, slot + 1);
  flush_sum(pushw, sh, slot + 2);
  flush_sum(pullw, sh, slot + 3);
  flush_max(maxa, sh, slot + 4);
  flush_max(maxi, sh, slot + 5);
  add_work(body.as.dg_work, ne, nv);
}

// ---------------------------------------------------------------- WM
template <class B>
__global__ void __launch_bounds__(BLOCK) k_pull_wm(B body) {
  using WA = typename B::WA;
  __shared__ int s_v[BLOCK];
  __shared__ int s_scan[BLOCK];
  __shared__ int s_flag[BLOCK];  // bit0: s_acc holds a message, bit1: done
  __shared__ int64_t s_start[BLOCK];
  __shared__ WA s_acc[BLOCK];
  const int lane = lane_id();
  const int off = threadIdx.x & ~(WAVE - 1);
  const int64_t N = body.as.get_size();
  const Status want = body.conf.want();
  const bool prune = body.conf.conf_pruning;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  unsigned long long ne = 0, nv = 0;
  for (int64_t base = wave << 6; base < N; base += nwaves << 6) {
    const int64_t idx = base + lane;
    const int v = (idx < N) ? body.as.fetch(idx, want) : -1;
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
    s_flag[off + lane] = 0;
    wave_sync();
    for (int j0 = 0; j0 < tot; j0 += WAVE) {
      const int j = j0 + lane;
      int o = -1;
      bool vote = false;
      WA m{};
      if (j < tot) {
        o = owner_search(&s_scan[off], WAVE, j);
        if (!(prune && (s_flag[off + o] & 2))) {
          const int64_t e = s_start[off + o] + (j - s_scan[off + o]);
          const int u = body.cols()[e];
          ++ne;
          if (body.source_ok(u)) {
            vote = true;
            m = body.msg(u, e);
          }
        }
      }
      body.seg_reduce(o, vote, m);
      if (body.seg_head(o) && vote) {
        const int k = off + o;
        if (s_flag[k] & 1)
          body.f.comp(&s_acc[k], m, body.g);
        else
          s_acc[k] = m;
        s_flag[k] |= prune ? 3 : 1;
      }
      wave_sync();
      if (prune) {
        // every target of this wavefront already found a parent
        const bool open = (v >= 0) && deg > 0 && !(s_flag[off + lane] & 2);
        if (ballot(open) == 0) break;
      }
    }
    if (v >= 0 && (s_flag[off + lane] & 1)) {
      WA acc = *body.f.wa_of(v);
      body.f.comp(&acc, s_acc[off + lane], body.g);
      *body.f.wa_of(v) = acc;
    }
    wave_sync();
  }
  add_work(body.as.dg_work, ne, nv);
}

// ---------------------------------------------------------------- CM
template <class B>
__global__ void __launch_bounds__(BLOCK) k_pull_cm(B body) {
  using WA = typename B::WA;
  __shared__ int s_v[BLOCK];
  __shared__ int s_scan[BLOCK];
  __shared__ int s_done[BLOCK];
  __shared__ int64_t s_start[BLOCK];
  __shared__ int s_red[BLOCK / WAVE];
  const int64_t N = body.as.get_size();
  const Status want = body.conf.want();
  const bool prune = body.conf.conf_pruning;
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
    s_done[threadIdx.x] = 0;
    __syncthreads();
    for (int j0 = 0; j0 < tot; j0 += BLOCK) {
      const int j = j0 + threadIdx.x;
      int o = -1;
      bool vote = false;
      WA m{};
      if (j < tot) {
        o = owner_search(s_scan, BLOCK, j);
        if (!(prune && s_done[o])) {
          const int64_t e = s_start[o] + (j - s_scan[o]);
          const int u = body.cols()[e];
          ++ne;
          if (body.source_ok(u)) {
            vote = true;
            m = body.msg(u, e);
          }
        }
      }
      body.seg_reduce(o, vote, m);
      if (body.seg_head(o) && vote) {
        body.f.compAtomic(body.f.wa_of(s_v[o]), m, body.g);
        if (prune) s_done[o] = 1;
      }
      __syncthreads();
    }
  }
  add_work(body.as.dg_work, ne, nv);
}

// ---------------------------------------------------------------- STRICT
// A segment that holds every in-edge of its target inside this wavefront's
// slot range is combined with a plain comp; only targets that straddle
// wavefronts (or workgroups) pay compAtomic.
// Workgroup `bi` of `nb` processes its equal share of the slots of `sl`
// (adjacency `cols`). `allow_whole` = false when other workgroups' slots may
// also feed the same target (TWOD slices), so every flush is compAtomic.
template <class B>
__device__ __forceinline__ void pull_strict_range(B &body, const slots_t sl,
                                                  const int32_t *cols,
                                                  const int32_t *tgt, int64_t N,
                                                  int64_t bi, int64_t nb,
                                                  bool allow_whole) {
  using WA = typename B::WA;
  __shared__ int64_t s_scan[SCHUNK + 1];
  __shared__ int64_t s_start[SCHUNK + 1];
  __shared__ int s_v[SCHUNK + 1];
  const Status want = body.conf.want();
  const bool prune = body.conf.conf_pruning;
  const int64_t total = sl.at(N);
  const int64_t per = (total + nb - 1) / nb;
  const int64_t lo = bi * per;
  const int64_t hi = MIN(total, lo + per);
  unsigned long long ne = 0;
  if (lo >= hi) return;
  const int wofs = threadIdx.x & ~(WAVE - 1);
  int64_t k0 = sl.last_le(N + 1, lo);
  for (int64_t c = lo; c < hi; c += SCHUNK) {
    strict_window(body, sl, N, k0, want, s_scan, s_v, s_start, tgt);
    __syncthreads();
    int o[SITEMS], u[SITEMS];
    int64_t e[SITEMS];
#pragma unroll
    for (int j = 0; j < SITEMS; ++j) {
      const int64_t slot = c + j * BLOCK + threadIdx.x;
      o[j] = -1;
      u[j] = -1;
      if (slot < hi) {
        o[j] = owner_search(s_scan, SCHUNK + 1, slot);
        const int64_t r = slot - s_scan[o[j]];
        const int64_t deg = s_scan[o[j] + 1] - s_scan[o[j]] - 1;
        const int v = s_v[o[j]];
        if (v >= 0 && r < deg && !(prune && body.as.bitmap.visited.test(v))) {
          e[j] = s_start[o[j]] + r;
          u[j] = 0;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < SITEMS; ++j)
      if (u[j] == 0) u[j] = cols[e[j]];
      else u[j] = -1;
#pragma unroll
    for (int j = 0; j < SITEMS; ++j) {
      bool vote = false;
      WA m{};
      if (u[j] >= 0) {
        ++ne;
        if (body.source_ok(u[j])) {
          vote = true;
          m = body.msg(u[j], e[j]);
        }
      }
      body.seg_reduce(o[j], vote, m);
      if (body.seg_head(o[j]) && vote) {
        const int v = s_v[o[j]];
        // this wavefront's slots for item j: [wlo, min(wlo + 64, hi))
        const int64_t wlo = c + j * BLOCK + wofs;
        const int64_t whi = MIN(wlo + WAVE, hi);
        const bool whole = allow_whole && s_scan[o[j]] >= wlo &&
                           s_scan[o[j] + 1] - 1 <= whi;
        if (prune) {
          if (body.as.bitmap.visited.test_and_set(v))
            body.f.compAtomic(body.f.wa_of(v), m, body.g);
        } else if (whole) {
          WA acc = *body.f.wa_of(v);
          body.f.comp(&acc, m, body.g);
          *body.f.wa_of(v) = acc;
        } else {
          body.f.compAtomic(body.f.wa_of(v), m, body.g);
        }
      }
    }
    const int64_t nk = strict_next_owner(s_scan, k0, c + SCHUNK);
    __syncthreads();
    k0 = nk;
  }
  add_work(body.as.dg_work, ne, 0);
}

// ------------------------------------------------- merge-path from-all pull
// For from-all pulls (every in-edge contributes: PageRank-like) over all
// targets. The slot space is the implicit rows[k] + k (deg + 1 slots per
// target, the last one closes the row). A tile of MP_TILE slots is staged
// in three steps:
//  A. the MP_TILE + 1 row prefixes and the tile's contiguous edge range are
//     read coalesced; every message emit(col[e]) is gathered into LDS, so all
//     gathers of the tile are in flight at once;
//  B. lane t walks its MP_R consecutive slots: one LDS read and one comp per
//     edge; a row that starts and ends inside the lane's run is written with
//     a plain comp (when `exclusive`), other row pieces become a head (row
//     that started before the run) or a tail (row still open at its end);
//  C. a head collects the tails of the lanes before it that belong to the
//     same row; rows wholly inside the tile are written plainly, rows that
//     cross a tile boundary (or any row, when !exclusive) use compAtomic.
// `exclusive` = this launch is the only producer for each target.
constexpr int MP_R = 8;
constexpr int MP_TILE = BLOCK * MP_R;

template <class B>
__device__ __forceinline__ void pull_mp_range(B &body, const int64_t *rows,
                                              const int32_t *cols,
                                              const int32_t *tgt, int64_t N,
                                              int64_t bi, int64_t nb,
                                              bool exclusive) {
  using WA = typename B::WA;
  __shared__ int s_row[MP_TILE + 2];  // row prefix relative to the tile start
  __shared__ WA s_msg[MP_TILE];
  __shared__ WA s_tail[BLOCK];
  __shared__ int s_tail_row[BLOCK];
  __shared__ int64_t s_k0;
  const int t = threadIdx.x;
  const Status want = body.conf.want();
  const slots_t sl{nullptr, rows};
  const int64_t total = sl.at(N);
  const int64_t ntiles = (total + MP_TILE - 1) / MP_TILE;
  const int64_t tpb = (ntiles + nb - 1) / nb;
  const int64_t tile_lo = bi * tpb, tile_hi = MIN(ntiles, tile_lo + tpb);
  unsigned long long ne = 0;
  if (tile_lo >= tile_hi) return;
  if (t == 0) s_k0 = sl.last_le(N + 1, tile_lo * MP_TILE);
  __syncthreads();
  int64_t k0 = s_k0;
  for (int64_t tile = tile_lo; tile < tile_hi; ++tile) {
    const int64_t S = tile * MP_TILE;
    const int64_t S_end = MIN(S + MP_TILE, total);
    const int span = (int)(S_end - S);
    // prefixes beyond the tile are clamped to MP_TILE + 1: only comparisons
    // against slots <= MP_TILE are ever made
    for (int j = t; j <= MP_TILE + 1; j += BLOCK) {
      const int64_t k = k0 + j;
      s_row[j] = (int)MIN(sl.at(k < N ? k : N) - S, (int64_t)MP_TILE + 1);
    }
    __syncthreads();
    // Slot s of row k is edge s - k, so the tile's edges are the contiguous
    // range [S - k0, S_end - k_end), k_end = owner of slot S_end (N at the
    // end of the slot space).
    int64_t k_end = (span < s_row[MP_TILE])
                        ? k0 + owner_search(s_row, MP_TILE + 1, span)
                        : k0 + MP_TILE;
    if (S_end == total) k_end = N;
    const int64_t e_lo = S - k0;
    const int nedge = (int)((S_end - k_end) - e_lo);
    for (int i = t; i < nedge; i += BLOCK) {
      const int64_t e = e_lo + i;
      s_msg[i] = body.msg(cols[e], e);
    }
    ne += (nedge + BLOCK - 1 - t) / BLOCK;
    __syncthreads();
    // B: sequential walk over [t*R, t*R + R)
    const int r0 = t * MP_R, r1 = MIN(span, r0 + MP_R);
    int tail_row = -1;
    WA tail{};
    int head_row = -1;
    WA head{};
    if (r0 < r1) {
      int k = owner_search(s_row, MP_TILE + 1, r0);
      int rs = s_row[k], re = s_row[k + 1] - 1;  // row start, closing slot
      WA acc{};
      bool has = false;
      for (int i = r0; i < r1; ++i) {
        if (i == re) {
          const int64_t v = tgt ? (int64_t)tgt[k0 + k] : k0 + k;
          if (rs >= r0) {  // the whole row lies in this lane's run
            if (has && body.as.fetch(v, want) >= 0) {
              if (exclusive) {
                WA w = *body.f.wa_of((int)v);
                body.f.comp(&w, acc, body.g);
                *body.f.wa_of((int)v) = w;
              } else {
                body.f.compAtomic(body.f.wa_of((int)v), acc, body.g);
              }
            }
          } else {
            head_row = has ? k : -2 - k;  // -2-k: head without messages
            head = acc;
          }
          ++k;
          rs = s_row[k];
          re = s_row[k + 1] - 1;
          has = false;
        } else {
          const WA m = s_msg[i - k];  // edge (S + i) - (k0 + k), tile-relative
          if (has)
            body.f.comp(&acc, m, body.g);
          else
            acc = m;
          has = true;
        }
      }
      if (has) {  // row k is still open after the run
        tail_row = k;
        tail = acc;
      }
    }
    s_tail_row[t] = tail_row;
    s_tail[t] = tail;
    __syncthreads();
    // C: heads gather the tails before them
    auto flush = [&](int k, WA val, bool inside) {
      const int64_t v = tgt ? (int64_t)tgt[k0 + k] : k0 + k;
      if (body.as.fetch(v, want) < 0) return;
      if (exclusive && inside) {
        WA w = *body.f.wa_of((int)v);
        body.f.comp(&w, val, body.g);
        *body.f.wa_of((int)v) = w;
      } else {
        body.f.compAtomic(body.f.wa_of((int)v), val, body.g);
      }
    };
    if (head_row != -1) {
      const int k = head_row >= 0 ? head_row : -2 - head_row;
      bool has = head_row >= 0;
      WA val = head;
      for (int q = t - 1; q >= 0 && s_tail_row[q] == k; --q) {
        if (has)
          body.f.comp(&val, s_tail[q], body.g);
        else
          val = s_tail[q];
        has = true;
      }
      if (has) flush(k, val, s_row[k] >= 0);
    }
    // the lane holding the tile's last slot closes a row that runs on
    if (r0 < span && span <= r0 + MP_R && tail_row >= 0) {
      WA val = tail;
      for (int q = t - 1; q >= 0 && s_tail_row[q] == tail_row; --q)
        body.f.comp(&val, s_tail[q], body.g);
      flush(tail_row, val, false);
    }
    // owner of the next tile start
    const int64_t nk = (MP_TILE < s_row[MP_TILE])
                           ? k0 + owner_search(s_row, MP_TILE + 1, MP_TILE)
                           : k0 + MP_TILE;
    __syncthreads();
    k0 = nk;
  }
  add_work(body.as.dg_work, ne, 0);
}

template <class B>
__global__ void __launch_bounds__(BLOCK) k_pull_mp(B body) {
  pull_mp_range(body, body.rows(), body.cols(), nullptr, body.as.get_size(),
                (int64_t)blockIdx.x, (int64_t)gridDim.x, true);
}

// ---------------------------------------------------------------- binned
// Binned pull over one TWOD slice (twod_slice_t): groups of G lanes reduce
// one row each (lane j of a group takes edges j, j + G, ...; a butterfly
// over the group combines them); the slice launch owns every row it
// writes, so rows are updated with a plain load and store. Rows longer
// than the last bin come as TWOD_CHUNK-edge chunks, one wavefront each; a
// row's only chunk is written plainly, several are combined into the target
// with compAtomic. No LDS, no tile phases. Each
// lane issues TWOD_UNROLL column loads and then their TWOD_UNROLL message
// gathers before combining (addresses clamped into the row, so the loads
// are unconditional): the row walk is bounded by memory-level parallelism,
// not by one load-to-use latency per edge.
template <int G, class B>
__device__ __forceinline__ void bin_walk(B &body, const int32_t *col, int64_t e0,
                                         int64_t e1, int sub,
                                         typename B::WA &acc, bool &has) {
  using WA = typename B::WA;
  constexpr int U = TWOD_UNROLL;
  for (int64_t eb = e0 + sub; eb < e1; eb += (int64_t)G * U) {
    int u[U];
    int64_t ek[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int64_t e = eb + (int64_t)k * G;
      ek[k] = e < e1 ? e : e1 - 1;
      u[k] = col[ek[k]];
    }
    WA m[U];
#pragma unroll
    for (int k = 0; k < U; ++k) m[k] = body.msg(u[k], ek[k]);
#pragma unroll
    for (int k = 0; k < U; ++k) {
      if (eb + (int64_t)k * G < e1) {
        if (has)
          body.f.comp(&acc, m[k], body.g);
        else
          acc = m[k];
        has = true;
      }
    }
  }
}
template <class B>
__device__ __forceinline__ void bin_combine(B &body, typename B::WA &acc,
                                            bool &has, int group) {
  for (int d = group >> 1; d > 0; d >>= 1) {
    const typename B::WA o = shfl_xor_any(acc, d);
    const int oh = __shfl_xor((int)has, d, 64);
    if (oh) {
      if (has)
        body.f.comp(&acc, o, body.g);
      else
        acc = o;
      has = true;
    }
  }
}

template <int G, class B>
__device__ __forceinline__ unsigned long long pull_bin_rows(
    B &body, const twod_slice_t &sl, const int32_t *rows, int64_t nrows,
    int64_t bi, int64_t nb) {
  using WA = typename B::WA;
  constexpr int GPW = WAVE / G;  // groups per wavefront
  const int lane = lane_id(), sub = lane & (G - 1);
  const Status want = body.conf.want();
  const int64_t wave = (bi * BLOCK + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = (nb * BLOCK) >> LOG_WAVE;
  unsigned long long ne = 0;
  // uniform over the wavefront: the butterfly needs every lane
  for (int64_t i0 = wave * GPW; i0 < nrows; i0 += nwaves * GPW) {
    const int64_t i = i0 + lane / G;
    WA acc{};
    bool has = false;
    int r = -1;
    if (i < nrows) {
      r = rows[i];
      const int64_t e0 = sl.off[r], e1 = sl.off[r + 1];
      bin_walk<G>(body, sl.col, e0, e1, sub, acc, has);
      if (sub == 0) ne += (unsigned long long)(e1 - e0);
    }
    bin_combine(body, acc, has, G);
    if (sub == 0 && has) {
      const int v = sl.tgt[r];
      if (body.as.fetch(v, want) >= 0) {
        WA w = *body.f.wa_of(v);
        body.f.comp(&w, acc, body.g);
        *body.f.wa_of(v) = w;
      }
    }
  }
  return ne;
}

template <class B>
__device__ __forceinline__ unsigned long long pull_bin_chunks(
    B &body, const twod_slice_t &sl, int64_t bi, int64_t nb) {
  using WA = typename B::WA;
  const int lane = lane_id();
  const Status want = body.conf.want();
  const int64_t wave = (bi * BLOCK + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = (nb * BLOCK) >> LOG_WAVE;
  unsigned long long ne = 0;
  for (int64_t k = wave; k < sl.n_chunks; k += nwaves) {
    const int r = sl.chunk_row[k];
    const int64_t e0 = sl.chunk_first[k];
    const int64_t e1 = MIN(e0 + TWOD_CHUNK, sl.off[r + 1]);
    WA acc{};
    bool has = false;
    bin_walk<WAVE>(body, sl.col, e0, e1, lane, acc, has);
    bin_combine(body, acc, has, WAVE);
    if (lane == 0) {
      ne += (unsigned long long)(e1 - e0);
      const int v = sl.tgt[r];
      if (has && body.as.fetch(v, want) >= 0) {
        if (e0 == sl.off[r] && e1 == sl.off[r + 1]) {  // the row's only chunk
          WA w = *body.f.wa_of(v);
          body.f.comp(&w, acc, body.g);
          *body.f.wa_of(v) = w;
        } else {
          body.f.compAtomic(body.f.wa_of(v), acc, body.g);
        }
      }
    }
  }
  return ne;
}

// workgroups [0, blocks[0]) serve bin 0, the next blocks[1] bin 1, the
// last blocks[2] the chunks of the longer rows
template <class B>
__global__ void __launch_bounds__(BLOCK) k_pull_bins(B body, twod_slice_t sl) {
  static_assert(TWOD_BINS == 2, "two row bins and the chunks");
  int64_t b = blockIdx.x;
  unsigned long long ne = 0;
  if (b < sl.blocks[0]) {
    ne = pull_bin_rows<twod_bin_group(0)>(body, sl, sl.bin_rows[0], sl.bin_n[0],
                                          b, sl.blocks[0]);
  } else if ((b -= sl.blocks[0]) < sl.blocks[1]) {
    ne = pull_bin_rows<twod_bin_group(1)>(body, sl, sl.bin_rows[1], sl.bin_n[1],
                                          b, sl.blocks[1]);
  } else {
    b -= sl.blocks[1];
    ne = pull_bin_chunks(body, sl, b, sl.blocks[2]);
  }
  add_work(body.as.dg_work, ne, 0);
}

template <class B>
__global__ void __launch_bounds__(BLOCK)
    k_pull_strict(B body, const int64_t *__restrict__ scan) {
  pull_strict_range(body, slots_t{scan, body.rows()}, body.cols(), nullptr,
                    body.as.get_size(), (int64_t)blockIdx.x, (int64_t)gridDim.x,
                    true);
}

template <class B>
__global__ void __launch_bounds__(BLOCK) k_pull_twod(B body, twod_slice_t sl) {
  pull_strict_range(body, slots_t{nullptr, sl.off}, sl.col, sl.tgt, sl.nr,
                    (int64_t)blockIdx.x, (int64_t)gridDim.x, true);
}

}  // namespace gsw
