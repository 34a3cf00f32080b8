// Propagation-blocked from-all pull (layout: pb_t in data_structures/graph.h).
//
// A from-all pull (every vertex sends emit(u) along each out-edge, every
// target combines what arrives) gathers one message per edge at a random
// source. On MI355X that gather is bound by the L2 request rate (about one
// request per edge, whether the line then hits L2 or not). Propagation
// blocking (Beamer, Asanovic, Patterson, "Reducing PageRank Communication
// via Propagation Blocking", IPDPS 2017) turns both random sides into LDS
// accesses at the price of one write and one read of a per-edge message:
//   phase 1  per source chunk (S consecutive ids): the chunk's messages are
//            computed once into LDS; for every tile (b, c) of the chunk the
//            edges' 16-bit source offsets are streamed and each edge's
//            message is stored at the edge's slot (consecutive lanes,
//            consecutive slots);
//   phase 2  per target bin: the bin's slots are one contiguous range
//            (bin-major layout); messages and 16-bit target offsets are
//            streamed and combined with compAtomic into LDS accumulators
//            that start from the targets' WA (from zero when the fused
//            filter keeps WA at zero, conf.wa_zero), which are then written
//            back (the workgroup owns its targets). A target with more in-edges
//            than a piece holds is a bin of its own, split over several
//            workgroups that each reduce their part in registers and add it
//            with one compAtomic on the target.
// HBM bytes per expand: 2 (source offset) + 2 * sizeof(msg) (store, load) +
// 2 (target offset) per edge, plus the messages and WA per vertex; no
// random global access. Messages are stored as F::msg_t when the functor
// declares one (its emit() result must convert to it exactly), else as WA.
//
// The layout is built once per graph on the device: bins by a host pass
// over the in-degree offsets, tile counts and slots by two passes over the
// in-CSR (one wavefront per row, one atomic per run of equal tiles), the
// pieces on the host from the tile offsets.
#pragma once
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "../data_structures/graph.h"
#include "expand_pull.h"
#include "scan.h"

namespace gsw {

constexpr int PB_BLOCK = 1024;
constexpr int PB_LDS_BYTES = 64 << 10;  // per phase: messages / accumulators
constexpr int PB_QUADS = 2;             // aligned 4-slot quads a lane keeps in flight (phase 1)
constexpr int PB_GQUADS = 4;            // the same in phase 2
constexpr int PB_XCDS = 8;              // MI355X: workgroups dealt round-robin over 8 XCDs

// four consecutive slots, loaded and stored as one access
template <typename T> struct alignas(4 * sizeof(T)) quad_t {
  T v[4];
};

template <typename F, typename = void> struct pb_msg {
  using type = typename F::wa_t;
};
template <typename F> struct pb_msg<F, std::void_t<typename F::msg_t>> {
  using type = typename F::msg_t;
};

// a functor with filter_acc(v, acc, g): its filter given the accumulated
// value, for a WA kept at zero (the fused gather of conf.wa_zero)
template <typename F, typename = void> struct pb_has_filter_acc : std::false_type {};
template <typename F>
struct pb_has_filter_acc<F, std::void_t<decltype(&F::filter_acc)>> : std::true_type {};

// bin of target v: the last b with bin_lo[b] <= v
__device__ __forceinline__ int64_t pb_bin_of(const int32_t *bin_lo, int64_t B,
                                             int64_t v) {
  int64_t lo = 0, hi = B;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (bin_lo[mid] <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// One wavefront per in-CSR row v (edges sorted by source, so tiles are
// non-decreasing along the row): count (cursor == nullptr) adds each run of
// equal tiles to cnt[tile]; fill (cursor != nullptr) claims slots for a run
// with one atomic and writes the edges' offsets.
__global__ void __launch_bounds__(BLOCK)
    k_pb_tiles(const int64_t *__restrict__ rows, const int32_t *__restrict__ cols,
               int64_t n, const int32_t *__restrict__ bin_lo, int64_t B, int S,
               int64_t C, unsigned long long *cnt, int32_t *cursor,
               const int64_t *__restrict__ tile_off, uint16_t *src, uint16_t *dst) {
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  for (int64_t v = wave; v < n; v += nwaves) {
    const int64_t s = rows[v], e = rows[v + 1];
    if (s == e) continue;
    const int64_t b = pb_bin_of(bin_lo, B, v);
    const int32_t vlo = bin_lo[b];
    for (int64_t k0 = s; k0 < e; k0 += WAVE) {
      const int64_t k = k0 + lane;
      const bool ok = k < e;
      const int u = ok ? cols[k] : 0;
      const int64_t tile = ok ? b * C + u / S : -1;
      const int64_t prev = __shfl_up(tile, 1, 64);
      const bool head = ok && (lane == 0 || prev != tile);
      const uint64_t hm = ballot(head);
      const int nvalid = (int)MIN((int64_t)WAVE, e - k0);
      const uint64_t later = hm & ~((2ull << lane) - 1);
      const int run = head ? (later ? __builtin_ctzll(later) : nvalid) - lane : 0;
      if (!cursor) {
        if (head) atomicAdd(cnt + tile, (unsigned long long)run);
      } else {
        int base = 0;
        if (head) base = atomicAdd(cursor + tile, run);
        const int hl = 63 - __clzll((long long)(hm & ((2ull << lane) - 1)));
        base = __shfl(base, hl, 64);
        if (ok) {
          const int64_t pos = tile_off[tile] + base + (lane - hl);
          src[pos] = (uint16_t)(u - (u / S) * S);
          dst[pos] = (uint16_t)(v - vlo);
        }
      }
    }
  }
}

// The same count / fill with one workgroup per bin, for layouts of at most
// PB_BIN_C chunks: the bin's tile counters (count) or cursors and tile
// offsets (fill) live in LDS, so a run of edges of one tile costs one LDS
// atomic instead of a global one. The bin's in-edges are one contiguous
// range (its targets are consecutive rows), walked 64 edges per wavefront
// step whatever the row lengths (a one-target bin of a hub is split over
// every wavefront); a lane finds its row by binary search among the bin's
// rows from the row of the wavefront's first edge (empty rows included).
constexpr int PB_BIN_C = 8192;

__device__ __forceinline__ int64_t pb_row_of(const int64_t *__restrict__ rows,
                                             int64_t lo, int64_t hi, int64_t k) {
  // the last r in [lo, hi) with rows[r] <= k
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (rows[mid] <= k)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(PB_BLOCK)
    k_pb_tiles_bin(const int64_t *__restrict__ rows, const int32_t *__restrict__ cols,
                   const int32_t *__restrict__ bin_lo, int S, int64_t C,
                   unsigned long long *cnt, const int64_t *__restrict__ tile_off,
                   uint16_t *src, uint16_t *dst) {
  __shared__ int64_t s_off[PB_BIN_C];
  __shared__ int s_cur[PB_BIN_C];
  const int64_t b = blockIdx.x;
  const bool fill = tile_off != nullptr;
  const int64_t vlo = bin_lo[b], vhi = bin_lo[b + 1];
  for (int64_t c = threadIdx.x; c < C; c += PB_BLOCK) {
    s_cur[c] = 0;
    if (fill) s_off[c] = tile_off[b * C + c];
  }
  __syncthreads();
  const int lane = lane_id();
  const int wv = threadIdx.x >> LOG_WAVE;
  constexpr int NW = PB_BLOCK / WAVE;
  const int64_t E0 = rows[vlo], E1 = rows[vhi];
  for (int64_t k0 = E0 + (int64_t)wv * WAVE; k0 < E1; k0 += (int64_t)NW * WAVE) {
    const int64_t r0 = pb_row_of(rows, vlo, vhi, k0);
    const int64_t k = k0 + lane;
    const bool ok = k < E1;
    const int64_t v = ok ? pb_row_of(rows, r0, vhi, k) : vlo;
    const int u = ok ? cols[k] : 0;
    const int c = u / S;
    const int prev = __shfl_up(c, 1, 64);
    const bool head = ok && (lane == 0 || prev != c);
    const uint64_t hm = ballot(head);
    const int nvalid = (int)MIN((int64_t)WAVE, E1 - k0);
    const uint64_t later = hm & ~((2ull << lane) - 1);
    const int run = head ? (later ? __builtin_ctzll(later) : nvalid) - lane : 0;
    if (!fill) {
      if (head) atomicAdd(&s_cur[c], run);
    } else {
      int base = 0;
      if (head) base = atomicAdd(&s_cur[c], run);
      const int hl = 63 - __clzll((long long)(hm & ((2ull << lane) - 1)));
      base = __shfl(base, hl, 64);
      if (ok) {
        const int64_t pos = s_off[c] + base + (lane - hl);
        src[pos] = (uint16_t)(u - c * S);
        dst[pos] = (uint16_t)(v - vlo);
      }
    }
  }
  if (!fill) {
    __syncthreads();
    for (int64_t c = threadIdx.x; c < C; c += PB_BLOCK)
      cnt[b * C + c] = (unsigned long long)s_cur[c];
  }
}

// One step of a phase-1 tile walk: the tile's bounds, its unaligned head
// and tail slots (one per lane) and its first PB_QUADS * 64 aligned quads
// of 16-bit source offsets, loaded before they are needed.
struct pb_step_t {
  int64_t start = 0, end = 0, a0 = 0, a1 = 0, nq = 0;
  uint16_t hs = 0, ts = 0;
  quad_t<uint16_t> si[PB_QUADS] = {};
};

__device__ __forceinline__ int64_t readlane64(int64_t x, int j) {
  const int lo = readlane((int)(uint32_t)x, j);
  const int hi = readlane((int)(x >> 32), j);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ void pb_load_step(pb_step_t &st, const uint16_t *src,
                                             int64_t start, int64_t len, int lane) {
  st.start = start;
  st.end = start + len;
  st.a0 = MIN(st.end, (start + 3) & ~(int64_t)3);
  st.a1 = st.a0 + ((st.end - st.a0) & ~(int64_t)3);
  st.nq = (st.a1 - st.a0) >> 2;
  if (lane < st.a0 - start) st.hs = src[start + lane];
  if (lane < st.end - st.a1) st.ts = src[st.a1 + lane];
  const quad_t<uint16_t> *sq = (const quad_t<uint16_t> *)(src + st.a0);
#pragma unroll
  for (int j = 0; j < PB_QUADS; ++j) {
    const int64_t q = j * WAVE + lane;
    if (q < st.nq) st.si[j] = sq[q];
  }
}

// Phase 1: one workgroup per piece (chunk c, bins [b0, b1)). The chunk's
// messages are staged in LDS once; then every wavefront takes whole tiles
// of the piece in turn (tile b0 + w, b0 + w + waves, ...): the bounds of
// up to 64 of its tiles load at once (one per lane), and the source offsets
// of the next tile load while the current tile's messages are stored, so a
// wavefront keeps two tiles in flight. A lane moves aligned quads of slots:
// one 8-byte load of four 16-bit source offsets, four LDS reads, one 16-byte
// store of four messages (for 4-byte messages); the at most three slots
// before the first aligned quad and after the last one take a lane each.
// Tiles hold about m / (B * C) edges (about 330 on R-MAT scale 22), so a
// tile per wavefront wastes few lanes and no lane searches for its tile.
template <class Bd, typename MSG>
__global__ void __launch_bounds__(PB_BLOCK) k_pb_scatter(Bd body, pb_t pb) {
  constexpr int CAP = PB_LDS_BYTES / (int)sizeof(MSG);
  __shared__ MSG s_msg[CAP];
  const int32_t *pc = pb.p1 + 4 * (int64_t)blockIdx.x;
  const int c = pc[0], b0 = pc[1], b1 = pc[2];
  if (b0 >= b1) return;  // padding of the XCD-aware piece order
  const int64_t lo = (int64_t)c * pb.S;
  const int w = (int)MIN((int64_t)pb.S, body.g.nvertexs - lo);
  MSG *vals = (MSG *)pb.vals;
  const int lane = lane_id();
  const int wv = threadIdx.x >> LOG_WAVE;
  constexpr int NW = PB_BLOCK / WAVE;
  const int64_t row = (int64_t)c * pb.B;
  const int mine = b1 - b0 > wv ? (b1 - b0 - wv + NW - 1) / NW : 0;
  // the first tiles' bounds and source offsets load before the messages
  // are staged
  int64_t my_start = 0, my_len = 0;
  pb_step_t cur, nxt;
  auto load_batch = [&](int t0) {
    const int cnt = MIN(WAVE, mine - t0);
    if (lane < cnt) {
      const int64_t t = row + b0 + wv + (int64_t)(t0 + lane) * NW;
      my_start = pb.cm_start[t];
      my_len = pb.cm_len[t];
    }
    pb_load_step(cur, pb.src, readlane64(my_start, 0), readlane64(my_len, 0), lane);
  };
  if (mine > 0) load_batch(0);
  for (int i = threadIdx.x; i < w; i += PB_BLOCK)
    s_msg[i] = (MSG)body.msg((int)(lo + i), 0);
  __syncthreads();
  for (int t0 = 0; t0 < mine; t0 += WAVE) {
    const int cnt = MIN(WAVE, mine - t0);
    if (t0 > 0) load_batch(t0);
    for (int j = 0; j < cnt; ++j) {
      if (j + 1 < cnt)
        pb_load_step(nxt, pb.src, readlane64(my_start, j + 1),
                     readlane64(my_len, j + 1), lane);
      if (lane < cur.a0 - cur.start) vals[cur.start + lane] = s_msg[cur.hs];
      if (lane < cur.end - cur.a1) vals[cur.a1 + lane] = s_msg[cur.ts];
      quad_t<MSG> *vq = (quad_t<MSG> *)(vals + cur.a0);
#pragma unroll
      for (int k = 0; k < PB_QUADS; ++k) {
        const int64_t q = k * WAVE + lane;
        if (q < cur.nq) {
          quad_t<MSG> o;
#pragma unroll
          for (int r = 0; r < 4; ++r) o.v[r] = s_msg[cur.si[k].v[r]];
          vq[q] = o;
        }
      }
      // the rest of a long tile
      const quad_t<uint16_t> *sq = (const quad_t<uint16_t> *)(pb.src + cur.a0);
      for (int64_t q0 = WAVE * PB_QUADS; q0 < cur.nq; q0 += WAVE * PB_QUADS) {
        quad_t<uint16_t> si[PB_QUADS] = {};
#pragma unroll
        for (int k = 0; k < PB_QUADS; ++k) {
          const int64_t q = q0 + k * WAVE + lane;
          if (q < cur.nq) si[k] = sq[q];
        }
#pragma unroll
        for (int k = 0; k < PB_QUADS; ++k) {
          const int64_t q = q0 + k * WAVE + lane;
          if (q < cur.nq) {
            quad_t<MSG> o;
#pragma unroll
            for (int r = 0; r < 4; ++r) o.v[r] = s_msg[si[k].v[r]];
            vq[q] = o;
          }
        }
      }
      cur = nxt;
    }
  }
}

// Phase 2: one workgroup per piece (bin b, slots [e0, e1), whole bin or
// part of a one-target bin). Slots are combined with compAtomic into LDS
// accumulators (native ds_add_f64 for PageRank), after the slots of one
// target within an aligned quad are combined in registers. (A segmented
// shuffle reduction over 64-slot groups cost more in shuffles than it saved
// in atomics.)
template <class Bd, typename MSG>
__global__ void __launch_bounds__(PB_BLOCK) k_pb_gather(Bd body, pb_t pb) {
  using WA = typename Bd::WA;
  constexpr int CAP = PB_LDS_BYTES / (int)sizeof(WA);
  __shared__ WA s_acc[CAP];
  __shared__ int s_has[PB_BLOCK / WAVE];
  const int64_t *q = pb.p2 + 4 * (int64_t)blockIdx.x;
  const int64_t b = q[0], e0 = q[1], e1 = q[2];
  const bool whole = q[3] != 0;
  const int32_t lo = pb.bin_lo[b];
  const int w = pb.bin_lo[b + 1] - lo;
  const Status want = body.conf.want();
  const MSG *vals = (const MSG *)pb.vals;
  const int lane = lane_id();
  if (whole) {
    // aligned quads of slots (one 16-byte message load and one 8-byte
    // offset load per quad for 4-byte messages), the next step's quads
    // loading while the current step's are combined (the first step's
    // before the accumulators are initialised); the at most three slots
    // before the first aligned quad and after the last one take a thread each
    const int t = threadIdx.x;
    const int64_t a0 = MIN(e1, (e0 + 3) & ~(int64_t)3);
    const int64_t a1 = a0 + ((e1 - a0) & ~(int64_t)3);
    const int64_t nq = (a1 - a0) >> 2;
    const quad_t<MSG> *vq = (const quad_t<MSG> *)(vals + a0);
    const quad_t<uint16_t> *dq = (const quad_t<uint16_t> *)(pb.dst + a0);
    constexpr int64_t STEP = (int64_t)PB_BLOCK * PB_GQUADS;
    quad_t<MSG> mv[PB_GQUADS], mn[PB_GQUADS];
    quad_t<uint16_t> di[PB_GQUADS], dn[PB_GQUADS];
    auto load = [&](int64_t q0, quad_t<MSG> *m, quad_t<uint16_t> *d) {
#pragma unroll
      for (int j = 0; j < PB_GQUADS; ++j) {
        const int64_t q = q0 + (int64_t)j * PB_BLOCK + t;
        const int64_t qq = q < nq ? q : nq - 1;
        m[j] = vq[qq];
        d[j] = dq[qq];
      }
    };
    if (nq > 0) load(0, mv, di);
    if (body.conf.pb_fuse_filter && body.conf.wa_zero) {
      for (int i = threadIdx.x; i < w; i += PB_BLOCK) s_acc[i] = WA{};
    } else {
      for (int i = threadIdx.x; i < w; i += PB_BLOCK) s_acc[i] = *body.f.wa_of(lo + i);
    }
    __syncthreads();
    if (t < a0 - e0) body.f.compAtomic(&s_acc[pb.dst[e0 + t]], (WA)vals[e0 + t], body.g);
    if (t < e1 - a1) body.f.compAtomic(&s_acc[pb.dst[a1 + t]], (WA)vals[a1 + t], body.g);
    for (int64_t q0 = 0; q0 < nq; q0 += STEP) {
      if (q0 + STEP < nq) load(q0 + STEP, mn, dn);
#pragma unroll
      for (int j = 0; j < PB_GQUADS; ++j)
        if (q0 + (int64_t)j * PB_BLOCK + t < nq) {
          // slots of one target are mostly consecutive (a target's in-edges
          // from one source chunk form a run), so a quad's slots of one
          // target combine in registers and cost one LDS atomic
          WA acc = (WA)mv[j].v[0];
          int d = di[j].v[0];
#pragma unroll
          for (int r = 1; r < 4; ++r) {
            if (di[j].v[r] == d) {
              body.f.comp(&acc, (WA)mv[j].v[r], body.g);
            } else {
              body.f.compAtomic(&s_acc[d], acc, body.g);
              d = di[j].v[r];
              acc = (WA)mv[j].v[r];
            }
          }
          body.f.compAtomic(&s_acc[d], acc, body.g);
        }
#pragma unroll
      for (int j = 0; j < PB_GQUADS; ++j) {
        mv[j] = mn[j];
        di[j] = dn[j];
      }
    }
    __syncthreads();
    bool applied = false;
    if constexpr (pb_has_filter_acc<decltype(body.f)>::value) {
      if (body.conf.pb_fuse_filter && body.conf.wa_zero) {
        // WA stays at zero: the filter takes the accumulator from LDS
        for (int i = threadIdx.x; i < w; i += PB_BLOCK) {
          const int v = lo + i;
          const WA a = body.as.fetch(v, want) >= 0 ? s_acc[i] : *body.f.wa_of(v);
          (void)body.f.filter_acc(v, a, body.g);
        }
        applied = true;
      }
    }
    if (!applied)
      for (int i = threadIdx.x; i < w; i += PB_BLOCK) {
        const int v = lo + i;
        if (body.as.fetch(v, want) >= 0) *body.f.wa_of(v) = s_acc[i];
        if (body.conf.pb_fuse_filter) (void)body.f.filter(v, body.g);
      }
  } else {
    // one target: reduce the part in registers, then across the workgroup
    WA acc{};
    bool has = false;
    for (int64_t k = e0 + threadIdx.x; k < e1; k += PB_BLOCK) {
      const WA m = (WA)vals[k];
      if (has)
        body.f.comp(&acc, m, body.g);
      else
        acc = m;
      has = true;
    }
    bin_combine(body, acc, has, WAVE);
    const int wv = threadIdx.x >> LOG_WAVE;
    if (lane == 0) {
      s_acc[wv] = acc;
      s_has[wv] = has;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      bool any = false;
      WA tot{};
      for (int i = 0; i < PB_BLOCK / WAVE; ++i) {
        if (!s_has[i]) continue;
        if (any)
          body.f.comp(&tot, s_acc[i], body.g);
        else
          tot = s_acc[i];
        any = true;
      }
      if (any && body.as.fetch(lo, want) >= 0)
        body.f.compAtomic(body.f.wa_of(lo), tot, body.g);
    }
  }
  add_work<PB_BLOCK>(body.as.dg_work, threadIdx.x == 0 ? (unsigned long long)(e1 - e0) : 0ull,
                     0ull);
}

// The functor's filter on the targets of split bins (pb_fuse_filter).
template <class Bd>
__global__ void __launch_bounds__(BLOCK)
    k_pb_filter_split(Bd body, const int32_t *__restrict__ split, int64_t k) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) (void)body.f.filter(split[i], body.g);
}

// Build the layout of `g` for chunks of S sources and bins of at most D
// targets (no-op if built with the same widths).
template <typename G> void pb_ensure(G &g, int S, int D) {
  GSW_REQUIRE(g.pb, E_INTERNAL_CHECK, "graph has no PB handle");
  GSW_REQUIRE(S >= 1 && S <= 65536 && D >= 1 && D <= 65536, E_INTERNAL_CHECK,
              "PB widths must fit 16-bit offsets");
  pb_t &pb = *g.pb;
  if (pb.src && pb.S == S && pb.D == D) return;
  pb.release();
  const int64_t n = g.nvertexs, m = g.nedges;
  GSW_REQUIRE(n >= 1, E_INVALID, "empty graph");
  // in-degree offsets on the host: bins of <= D targets and <= ecap in-edges
  std::vector<int64_t> rp((size_t)n + 1);
  H_ERR(hipMemcpy(rp.data(), g.dgr_row_ptr, sizeof(int64_t) * (n + 1),
                  hipMemcpyDeviceToHost));
  int64_t ecap = MAX((int64_t)16384, m / 1024);
  if (const char *e = std::getenv("GSW_PB_ECAP")) {
    const long long x = std::atoll(e);
    if (x > 0) ecap = x;
  }
  std::vector<int32_t> blo(1, 0);
  for (int64_t v = 0; v < n;) {
    int64_t w = v + 1;
    while (w < n && w - v < D && rp[w + 1] - rp[v] <= ecap) ++w;
    blo.push_back((int32_t)w);
    v = w;
  }
  pb.S = S;
  pb.D = D;
  pb.m = m;
  pb.ecap = ecap;
  pb.B = (int64_t)blo.size() - 1;
  pb.C = (n + S - 1) / S;
  const int64_t T = pb.B * pb.C;
  H_ERR(hipMalloc((void **)&pb.bin_lo, sizeof(int32_t) * (pb.B + 1)));
  H_ERR(hipMemcpy(pb.bin_lo, blo.data(), sizeof(int32_t) * (pb.B + 1),
                  hipMemcpyHostToDevice));
  H_ERR(hipMalloc((void **)&pb.tile_off, sizeof(int64_t) * (T + 1)));
  H_ERR(hipMalloc((void **)&pb.src, sizeof(uint16_t) * (m + 1)));
  H_ERR(hipMalloc((void **)&pb.dst, sizeof(uint16_t) * (m + 1)));
  int32_t *cursor = nullptr;
  int64_t *bsum = nullptr;
  H_ERR(hipMalloc((void **)&cursor, sizeof(int32_t) * (T + 1)));
  H_ERR(hipMalloc((void **)&bsum, sizeof(int64_t) * ((T + 1 + SCAN_TILE - 1) / SCAN_TILE + 2)));
  fill_async<int64_t>(pb.tile_off, 0, T + 1);
  fill_async<int32_t>(cursor, 0, T + 1);
  const int grid = grid_for(n * WAVE);
  // (int LDS counters: a bin holds at most m edges)
  const bool per_bin = pb.C <= PB_BIN_C && pb.B <= INT32_MAX && m <= INT32_MAX;
  if (per_bin) {
    hipLaunchKernelGGL(k_pb_tiles_bin, dim3((unsigned)pb.B), dim3(PB_BLOCK), 0, stream(),
                       (const int64_t *)g.dgr_row_ptr, (const int32_t *)g.dgr_adj_list,
                       (const int32_t *)pb.bin_lo, S, pb.C,
                       (unsigned long long *)pb.tile_off, nullptr, nullptr, nullptr);
  } else {
    hipLaunchKernelGGL(k_pb_tiles, dim3(grid), dim3(BLOCK), 0, stream(),
                       (const int64_t *)g.dgr_row_ptr, (const int32_t *)g.dgr_adj_list,
                       n, (const int32_t *)pb.bin_lo, pb.B, S, pb.C,
                       (unsigned long long *)pb.tile_off, nullptr, nullptr, nullptr,
                       nullptr);
  }
  H_ERR(hipGetLastError());
  scan_exclusive_inplace(pb.tile_off, T, bsum);
  if (per_bin) {
    hipLaunchKernelGGL(k_pb_tiles_bin, dim3((unsigned)pb.B), dim3(PB_BLOCK), 0, stream(),
                       (const int64_t *)g.dgr_row_ptr, (const int32_t *)g.dgr_adj_list,
                       (const int32_t *)pb.bin_lo, S, pb.C, nullptr,
                       (const int64_t *)pb.tile_off, pb.src, pb.dst);
  } else {
    hipLaunchKernelGGL(k_pb_tiles, dim3(grid), dim3(BLOCK), 0, stream(),
                       (const int64_t *)g.dgr_row_ptr, (const int32_t *)g.dgr_adj_list,
                       n, (const int32_t *)pb.bin_lo, pb.B, S, pb.C, nullptr, cursor,
                       (const int64_t *)pb.tile_off, pb.src, pb.dst);
  }
  H_ERR(hipGetLastError());
  std::vector<int64_t> off((size_t)T + 1);
  H_ERR(hipMemcpyAsync(off.data(), pb.tile_off, sizeof(int64_t) * (T + 1),
                       hipMemcpyDeviceToHost, stream()));
  sync();
  hipFree(cursor);
  hipFree(bsum);
  GSW_REQUIRE(off[T] == m, E_INTERNAL_CHECK, "PB tiles do not cover the edges");
  // groups: runs of consecutive bins of about m / K edges (GSW_PB_GROUPS = K,
  // default 1); an expand runs phase 1 then phase 2 group by group, so a
  // group's phase 2 reads back messages its phase 1 has just written
  int64_t K = 1;
  if (const char *e = std::getenv("GSW_PB_GROUPS")) {
    const long long x = std::atoll(e);
    if (x > 1) K = MIN((int64_t)x, pb.B);
  }
  std::vector<int32_t> grp((size_t)pb.B);
  {
    int64_t k = 0;
    for (int64_t b = 0; b < pb.B; ++b) {
      // bin b starts a new group once the edges before it reach k+1 shares
      while (k + 1 < K && off[b * pb.C] * K >= (k + 1) * m) ++k;
      grp[b] = (int32_t)k;
    }
  }
  // chunk-major tile table and phase-1 pieces (runs of bins of <= ecap1
  // edges within one group; ecap1 = ecap unless GSW_PB_ECAP1 is set),
  // listed group by group
  int64_t ecap1 = ecap;
  if (const char *e = std::getenv("GSW_PB_ECAP1")) {
    const long long x = std::atoll(e);
    if (x > 0) ecap1 = x;
  }
  std::vector<int64_t> cs((size_t)T);
  std::vector<int32_t> cl((size_t)T);
  std::vector<std::vector<int32_t>> p1g((size_t)K);
  for (int64_t c = 0; c < pb.C; ++c) {
    int64_t acc = 0, b0 = 0;
    for (int64_t b = 0; b < pb.B; ++b) {
      const int64_t t = b * pb.C + c;
      const int64_t len = off[t + 1] - off[t];
      cs[c * pb.B + b] = off[t];
      cl[c * pb.B + b] = (int32_t)len;
      if (acc > 0 && (acc + len > ecap1 || grp[b] != grp[b0])) {
        p1g[grp[b0]].insert(p1g[grp[b0]].end(), {(int32_t)c, (int32_t)b0, (int32_t)b, 0});
        b0 = b;
        acc = 0;
      }
      if (acc == 0) b0 = b;
      acc += len;
    }
    if (acc > 0)
      p1g[grp[b0]].insert(p1g[grp[b0]].end(), {(int32_t)c, (int32_t)b0, (int32_t)pb.B, 0});
  }
  // XCD-aware order: workgroup i of a launch runs on XCD i % PB_XCDS, so a
  // group's pieces are dealt out such that every XCD takes one contiguous
  // run of the chunk-ordered list. The pieces of one chunk then run side by
  // side on one XCD and restage the chunk's messages from its L2. Slots
  // past the end are empty pieces (b0 == b1), which return at once.
  // GSW_PB_XCD=0 keeps the chunk order.
  bool deal = true;
  if (const char *e = std::getenv("GSW_PB_XCD")) deal = std::atoi(e) != 0;
  std::vector<int32_t> p1;
  std::vector<int64_t> g1(1, 0);
  for (int64_t k = 0; k < K; ++k) {
    const int64_t np = (int64_t)p1g[k].size() / 4;
    if (deal && np > PB_XCDS) {
      const int64_t per = (np + PB_XCDS - 1) / PB_XCDS;
      std::vector<int32_t> o((size_t)(4 * per * PB_XCDS), 0);
      for (int64_t b = 0; b < per * PB_XCDS; ++b) {
        const int64_t i = (b % PB_XCDS) * per + b / PB_XCDS;
        if (i < np)
          for (int r = 0; r < 4; ++r) o[4 * b + r] = p1g[k][4 * i + r];
      }
      p1g[k].swap(o);
    }
    p1.insert(p1.end(), p1g[k].begin(), p1g[k].end());
    g1.push_back((int64_t)p1.size() / 4);
  }
  // phase-2 pieces: whole bins, or ecap-slot parts of a one-target bin
  std::vector<int64_t> p2;
  std::vector<int32_t> split;
  for (int64_t b = 0; b < pb.B; ++b) {
    const int64_t e0 = off[b * pb.C], e1 = off[(b + 1) * pb.C];
    if (blo[b + 1] - blo[b] == 1 && e1 - e0 > ecap) {
      split.push_back(blo[b]);
      for (int64_t k = e0; k < e1; k += ecap)
        p2.insert(p2.end(), {b, k, MIN(e1, k + ecap), (int64_t)0});
    } else {
      p2.insert(p2.end(), {b, e0, e1, (int64_t)1});
    }
  }
  pb.n1 = (int64_t)p1.size() / 4;
  pb.n2 = (int64_t)p2.size() / 4;
  // phase-2 pieces are in bin order: group k starts at its first bin's piece
  std::vector<int64_t> g2((size_t)K + 1);
  for (int64_t k = 0, i = 0; k <= K; ++k) {
    while (i < pb.n2 && grp[p2[4 * i]] < k) ++i;
    g2[k] = i;
  }
  GSW_REQUIRE((int64_t)g1.size() == K + 1 && (int64_t)g2.size() == K + 1,
              E_INTERNAL_CHECK, "PB groups");
  pb.ng = K;
  pb.g1h = new int64_t[K + 1];
  pb.g2h = new int64_t[K + 1];
  for (int64_t k = 0; k <= K; ++k) {
    pb.g1h[k] = g1[k];
    pb.g2h[k] = g2[k];
  }
  // the pieces cover every edge once: phase 1 by chunk, phase 2 in order
  int64_t covered = 0;
  for (int64_t i = 0; i < pb.n1; ++i)
    for (int64_t b = p1[4 * i + 1]; b < p1[4 * i + 2]; ++b)
      covered += cl[(int64_t)p1[4 * i] * pb.B + b];
  GSW_REQUIRE(covered == m, E_INTERNAL_CHECK, "PB phase-1 pieces miss edges");
  for (int64_t i = 0, next = 0; i < pb.n2; ++i) {
    GSW_REQUIRE(p2[4 * i + 1] == next && p2[4 * i + 2] >= next &&
                    blo[p2[4 * i] + 1] - blo[p2[4 * i]] <= D,
                E_INTERNAL_CHECK, "PB phase-2 pieces out of order");
    next = p2[4 * i + 2];
    if (i + 1 == pb.n2)
      GSW_REQUIRE(next == m, E_INTERNAL_CHECK, "PB phase-2 pieces miss edges");
  }
  H_ERR(hipMalloc((void **)&pb.cm_start, sizeof(int64_t) * (T + 1)));
  H_ERR(hipMalloc((void **)&pb.cm_len, sizeof(int32_t) * (T + 1)));
  H_ERR(hipMalloc((void **)&pb.p1, sizeof(int32_t) * (p1.size() + 4)));
  H_ERR(hipMalloc((void **)&pb.p2, sizeof(int64_t) * (p2.size() + 4)));
  H_ERR(hipMemcpy(pb.cm_start, cs.data(), sizeof(int64_t) * T, hipMemcpyHostToDevice));
  H_ERR(hipMemcpy(pb.cm_len, cl.data(), sizeof(int32_t) * T, hipMemcpyHostToDevice));
  if (!p1.empty())
    H_ERR(hipMemcpy(pb.p1, p1.data(), sizeof(int32_t) * p1.size(), hipMemcpyHostToDevice));
  H_ERR(hipMemcpy(pb.p2, p2.data(), sizeof(int64_t) * p2.size(), hipMemcpyHostToDevice));
  pb.nsplit = (int64_t)split.size();
  H_ERR(hipMalloc((void **)&pb.split, sizeof(int32_t) * (split.size() + 1)));
  if (!split.empty())
    H_ERR(hipMemcpy(pb.split, split.data(), sizeof(int32_t) * split.size(),
                    hipMemcpyHostToDevice));
}

// Message storage of `bytes` per edge (grown on demand).
inline void pb_vals(pb_t &pb, size_t bytes) {
  const size_t need = bytes * (size_t)(pb.m + 1);
  if (pb.vals_bytes >= need) return;
  if (pb.vals) H_ERR(hipFree(pb.vals));
  pb.vals = nullptr;
  H_ERR(hipMalloc(&pb.vals, need));
  pb.vals_bytes = need;
}

}  // namespace gsw
