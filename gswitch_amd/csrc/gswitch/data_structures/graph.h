// Graphs on the device.
//
// device_graph_t<CSR, E> keeps the reference's names and accessors
// (src/data_structures/graph.cuh:44-175: dg_adj_list, dg_edgedata,
// get_out_degree, get_in_degree, get_in_start_pos, get_level, update_level)
// but stores int64 offsets row_ptr[n+1] instead of int32 start_pos[n] +
// odegree[n] texture objects; degrees come from adjacent offsets.
// device_graph_t<COO, E> is the edge-centric form (graph.cuh:12-41): one
// packed (v0 << 32 | v1) per undirected edge.
#pragma once
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "../../host/host_graph.h"
#include "../utils/common.h"

typedef int64_t packed_t;

// Degree statistics used as selector features. Restates
// graph_attr_t::fill (src/utils/fileIO.cuh:35-67): mean, sample standard
// deviation, range and max of the out-degrees, Gini index GI and normalised
// degree entropy Her over the ascending-sorted degrees.
struct graph_attr_t {
  double avg_deg = 0, std_deg = 0, max_deg = 0, range = 0, GI = 0, Her = 0;

  void fill(const int64_t *row_ptr, int64_t nv) {
    if (nv <= 0) return;
    std::vector<double> d((size_t)nv);
    for (int64_t i = 0; i < nv; ++i) d[i] = (double)(row_ptr[i + 1] - row_ptr[i]);
    std::sort(d.begin(), d.end());
    double s = 0, s2 = 0;
    for (double x : d) {
      s += x;
      s2 += x * x;
    }
    std_deg = nv > 1 ? std::sqrt((s2 - s * s / nv) / (nv - 1)) : 0.0;
    avg_deg = s / nv;
    range = d[nv - 1] - d[0];
    max_deg = d[nv - 1];
    double gi = 0, her = 0;
    for (int64_t i = 1; i <= nv; ++i) {
      double du = d[i - 1];
      if (s > 0) gi += (nv - i + 0.5) * du / nv / s;
      if (du != 0) her += du * (std::log(s) - std::log(du)) / s;
    }
    GI = 1 - 2 * gi;
    Her = nv > 1 ? her / std::log((double)nv) : 0.0;
  }
};

// Source-sliced layout for from-all pulls, the TWOD variant. The
// in-adjacency is cut into P slices of consecutive source ids, P chosen so
// that one slice's source-side data (the message each source emits) spans
// about TWOD_SLICE_BYTES. Slice p keeps only its non-empty target rows
// (doubly compressed): tgt[r] is the target vertex of row r (ascending),
// off[r..r+1) its range in col, and col holds the sources. The slices are
// pulled one launch after another, so each launch is the only producer for
// its targets and rows are updated with plain loads and stores (cf.
// chunk_t, src/utils/2d_partition.cuh:13-299, which tiles for shared
// memory instead). On MI355X the pull is bound by the L2 request rate of
// its message gathers (about one request per edge), not by where the lines
// come from: slices of 2-16 MB time within 7% of each other on R-MAT scale
// 22, so the default slice is 16 MB (one slice up to 4 Mi vertices of FP32
// state); GSW_TWOD_SLICE_KB overrides it.
//
// For the binned pull the rows of a slice are also listed by length: bin c
// holds the rows of length in (twod_bin_len(c-1), twod_bin_len(c)], served
// by groups of twod_bin_group(c) lanes; longer rows are cut into chunks of
// TWOD_CHUNK edges, one wavefront each. Every group and every chunk is one
// unrolled step of TWOD_UNROLL edges per lane, so work items are uniform.
// `blocks[c]` (c = 0..TWOD_BINS-1 bins, TWOD_BINS chunks) is the workgroup
// count a launch gives each part, in proportion to its wavefront steps.
constexpr int TWOD_MAX = 64;
constexpr int64_t TWOD_SLICE_BYTES = 16ll << 20;
constexpr int TWOD_BINS = 2;
constexpr int TWOD_UNROLL = 8;  // edges a lane keeps in flight
__host__ __device__ constexpr int twod_bin_group(int c) { return c == 0 ? 1 : 8; }
__host__ __device__ constexpr int twod_bin_len(int c) {
  return twod_bin_group(c) * TWOD_UNROLL;
}
constexpr int64_t TWOD_CHUNK = 64 * TWOD_UNROLL;
struct twod_slice_t {
  int64_t nr = 0;          // non-empty target rows
  int64_t ne = 0;          // edges
  int32_t *tgt = nullptr;  // [nr]
  int64_t *off = nullptr;  // [nr + 1]
  int32_t *col = nullptr;  // [ne]
  int32_t *bin_rows[TWOD_BINS] = {};
  int64_t bin_n[TWOD_BINS] = {};
  int32_t *chunk_row = nullptr;    // [n_chunks]
  int64_t *chunk_first = nullptr;  // [n_chunks] first edge of the chunk
  int64_t n_chunks = 0;
  int32_t blocks[TWOD_BINS + 1] = {};
};
struct twod_t {
  int P = 0;
  int32_t bound[TWOD_MAX + 1] = {};  // slice p: sources [bound[p], bound[p+1])
  twod_slice_t s[TWOD_MAX];
};

// Propagation-blocked layout for from-all pulls (kernels/pb.h builds it on
// first use). Sources are cut into chunks of S consecutive ids, targets
// into bins of at most D consecutive ids and about ecap in-edges (a target
// with more in-edges is a bin of its own). Tile (b, c) holds the edges from
// chunk c into bin b; the edges are stored bin-major ([b][c] tiles, each
// contiguous), one 16-bit source offset (within its chunk) and one 16-bit
// target offset (within its bin) per edge. One expand is two launches:
// phase 1, one workgroup per piece of a chunk, stages the chunk's messages
// in LDS and writes every edge's message into `vals` at the edge's slot
// (coalesced stores, tile by tile); phase 2, one workgroup per bin (or per
// piece of a one-target bin), streams the bin's slots and combines them
// into LDS accumulators. Every random access is an LDS access.
struct pb_t {
  int S = 0, D = 0;              // chunk width, largest bin width
  int64_t C = 0, B = 0;          // chunks, bins
  int64_t m = 0, ecap = 0;       // edges, edges per piece
  int32_t *bin_lo = nullptr;     // [B + 1] first target of each bin
  uint16_t *src = nullptr;       // [m] source - chunk start (bin-major)
  uint16_t *dst = nullptr;       // [m] target - bin start
  int64_t *tile_off = nullptr;   // [B * C + 1] bin-major tile offsets
  int64_t *cm_start = nullptr;   // [C * B] tile (b, c) start, chunk-major
  int32_t *cm_len = nullptr;     // [C * B] its length
  int64_t n1 = 0;                // phase-1 pieces
  int32_t *p1 = nullptr;         // [n1][4]: chunk, first bin, end bin, 0
  int64_t n2 = 0;                // phase-2 pieces
  int64_t *p2 = nullptr;         // [n2][4]: bin, first slot, end slot, whole bin
  void *vals = nullptr;          // [m] messages (msg type of the functor)
  size_t vals_bytes = 0;
  int32_t *split = nullptr;      // [nsplit] targets of bins split over pieces
  int64_t nsplit = 0;
  int64_t ng = 1;                // groups of consecutive bins (host arrays:
  int64_t *g1h = nullptr;        //   [ng + 1] first phase-1 piece of each,
  int64_t *g2h = nullptr;        //   [ng + 1] first phase-2 piece of each)
  void release() {
    delete[] g1h;
    delete[] g2h;
    hipFree(bin_lo);
    hipFree(src);
    hipFree(dst);
    hipFree(tile_off);
    hipFree(cm_start);
    hipFree(cm_len);
    hipFree(p1);
    hipFree(p2);
    hipFree(vals);
    hipFree(split);
    *this = pb_t();
  }
};

template <GraphFmt fmt, typename E> struct device_graph_t {};

template <typename E> struct device_graph_t<CSR, E> {
  typedef E edge_t;
  static constexpr bool has_edata = !std::is_same<E, Empty>::value;

  int64_t nvertexs = 0;
  int64_t nedges = 0;  // directed CSR entries
  int level = 0;
  bool directed = false;

  int64_t *dg_row_ptr = nullptr;
  int32_t *dg_adj_list = nullptr;
  E *dg_edgedata = nullptr;
  // reversed graph (directed only; aliases the forward arrays otherwise)
  int64_t *dgr_row_ptr = nullptr;
  int32_t *dgr_adj_list = nullptr;
  E *dgr_edgedata = nullptr;
  // first in-neighbour of every vertex (dgr_adj_list[dgr_row_ptr[v]], -1
  // without in-edges): the bottom-up pull's first probe of a wavefront of
  // consecutive targets is one coalesced load instead of one scattered
  // adjacency load per lane
  int32_t *dgr_first = nullptr;
  // out- and in-degrees as int32 (the reference's dg_odegree,
  // graph.cuh:44-175): per-vertex passes that need only a degree read 4
  // bytes instead of an 8-byte offset pair; the in-degrees alias the
  // out-degrees of an undirected graph
  int32_t *dg_odegree = nullptr;
  int32_t *dgr_idegree = nullptr;
  // host-side handle of the TWOD slices (kernels/twod.h builds them on first
  // use); shared by every copy of this struct, owned by the one that built it
  twod_t *twod = nullptr;
  // host-side handle of the propagation-blocked layout (kernels/pb.h), owned
  // the same way
  pb_t *pb = nullptr;

  // Upload a host CSR. `w` is converted to E when E is not Empty.
  int64_t build(const gsw_host::HostGraph &hg) {
    nvertexs = hg.n;
    nedges = hg.m;
    directed = hg.directed;
    level = 0;
    twod = new twod_t();
    pb = new pb_t();
    int64_t bytes = 0;
    H_ERR(hipMalloc((void **)&dg_row_ptr, sizeof(int64_t) * (nvertexs + 1)));
    H_ERR(hipMemcpy(dg_row_ptr, hg.row_ptr.data(),
                    sizeof(int64_t) * (nvertexs + 1), hipMemcpyHostToDevice));
    H_ERR(hipMalloc((void **)&dg_adj_list, sizeof(int32_t) * (nedges + 1)));
    if (nedges)
      H_ERR(hipMemcpy(dg_adj_list, hg.col.data(), sizeof(int32_t) * nedges,
                      hipMemcpyHostToDevice));
    bytes += sizeof(int64_t) * (nvertexs + 1) + sizeof(int32_t) * nedges;
    if (has_edata) {
      GSW_REQUIRE(hg.weighted, gsw::E_INVALID, "graph has no edge weights");
      bytes += upload_w(hg.w, &dg_edgedata);
    }
    if (directed) {
      H_ERR(hipMalloc((void **)&dgr_row_ptr, sizeof(int64_t) * (nvertexs + 1)));
      H_ERR(hipMemcpy(dgr_row_ptr, hg.r_row_ptr.data(),
                      sizeof(int64_t) * (nvertexs + 1), hipMemcpyHostToDevice));
      H_ERR(hipMalloc((void **)&dgr_adj_list, sizeof(int32_t) * (nedges + 1)));
      if (nedges)
        H_ERR(hipMemcpy(dgr_adj_list, hg.r_col.data(), sizeof(int32_t) * nedges,
                        hipMemcpyHostToDevice));
      bytes += sizeof(int64_t) * (nvertexs + 1) + sizeof(int32_t) * nedges;
      if (has_edata) bytes += upload_w(hg.r_w, &dgr_edgedata);
    } else {
      dgr_row_ptr = dg_row_ptr;
      dgr_adj_list = dg_adj_list;
      dgr_edgedata = dg_edgedata;
    }
    {
      const std::vector<int64_t> &rp = directed ? hg.r_row_ptr : hg.row_ptr;
      const std::vector<int32_t> &cl = directed ? hg.r_col : hg.col;
      std::vector<int32_t> first((size_t)nvertexs + 1, -1);
      for (int64_t v = 0; v < nvertexs; ++v)
        if (rp[v] < rp[v + 1]) first[v] = cl[rp[v]];
      H_ERR(hipMalloc((void **)&dgr_first, sizeof(int32_t) * (nvertexs + 1)));
      H_ERR(hipMemcpy(dgr_first, first.data(), sizeof(int32_t) * (nvertexs + 1),
                      hipMemcpyHostToDevice));
      bytes += sizeof(int32_t) * (nvertexs + 1);
    }
    bytes += upload_degrees(hg.row_ptr, &dg_odegree);
    if (directed)
      bytes += upload_degrees(hg.r_row_ptr, &dgr_idegree);
    else
      dgr_idegree = dg_odegree;
    return bytes;
  }

  void release() {
    if (twod) {
      for (int p = 0; p < twod->P; ++p) {
        twod_slice_t &sl = twod->s[p];
        hipFree(sl.tgt);
        hipFree(sl.off);
        hipFree(sl.col);
        for (int c = 0; c < TWOD_BINS; ++c) hipFree(sl.bin_rows[c]);
        hipFree(sl.chunk_row);
        hipFree(sl.chunk_first);
      }
      delete twod;
      twod = nullptr;
    }
    if (pb) {
      pb->release();
      delete pb;
      pb = nullptr;
    }
    if (directed) {
      hipFree(dgr_row_ptr);
      hipFree(dgr_adj_list);
      if (has_edata) hipFree(dgr_edgedata);
    }
    hipFree(dg_row_ptr);
    hipFree(dg_adj_list);
    hipFree(dgr_first);
    if (directed) hipFree(dgr_idegree);
    hipFree(dg_odegree);
    if (has_edata) hipFree(dg_edgedata);
    dgr_first = nullptr;
    dg_odegree = dgr_idegree = nullptr;
    dg_row_ptr = dgr_row_ptr = nullptr;
    dg_adj_list = dgr_adj_list = nullptr;
    dg_edgedata = dgr_edgedata = nullptr;
  }

  __device__ __forceinline__ int64_t get_start_pos(int v) const {
    return dg_row_ptr[v];
  }
  __device__ __forceinline__ int get_out_degree(int v) const {
    return dg_odegree[v];
  }
  __device__ __forceinline__ int get_in_degree(int v) const {
    return dgr_idegree[v];
  }
  __device__ __forceinline__ int64_t get_in_start_pos(int v) const {
    return dgr_row_ptr[v];
  }
  __device__ __forceinline__ E *fetch_edata(int64_t eid) const {
    if constexpr (has_edata) return dg_edgedata + eid;
    else return nullptr;
  }
  __device__ __forceinline__ E *fetch_edata_r(int64_t eid) const {
    if constexpr (has_edata) return dgr_edgedata + eid;
    else return nullptr;
  }
  // The level lives in the struct (passed by value, as in the reference) or,
  // for captured level loops, in device memory that a kernel advances.
  __device__ __forceinline__ int get_level() const {
    return level_ptr ? *level_ptr : level;
  }
  __host__ __device__ inline void update_level(int inc = 1) { level += inc; }
  __host__ __device__ inline void reset_level() { level = 0; }
  int *level_ptr = nullptr;

 private:
  static int64_t upload_degrees(const std::vector<int64_t> &rp, int32_t **dst) {
    const size_t n = rp.empty() ? 0 : rp.size() - 1;
    std::vector<int32_t> d(n + 1, 0);
    for (size_t v = 0; v < n; ++v) {
      const int64_t x = rp[v + 1] - rp[v];
      GSW_REQUIRE(x <= INT32_MAX, gsw::E_INVALID, "a vertex degree exceeds int32");
      d[v] = (int32_t)x;
    }
    H_ERR(hipMalloc((void **)dst, sizeof(int32_t) * (n + 1)));
    H_ERR(hipMemcpy(*dst, d.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice));
    return (int64_t)(sizeof(int32_t) * n);
  }
  int64_t upload_w(const std::vector<uint32_t> &w, E **dst) {
    std::vector<E> tmp(w.size());
    for (size_t i = 0; i < w.size(); ++i) tmp[i] = (E)w[i];
    H_ERR(hipMalloc((void **)dst, sizeof(E) * (tmp.size() + 1)));
    if (!tmp.empty())
      H_ERR(hipMemcpy(*dst, tmp.data(), sizeof(E) * tmp.size(),
                      hipMemcpyHostToDevice));
    return (int64_t)(sizeof(E) * tmp.size());
  }
};

template <typename E> struct device_graph_t<COO, E> {
  typedef E edge_t;
  int64_t nvertexs = 0;
  int64_t nedges = 0;  // undirected edges, one packed pair each
  int level = 0;
  packed_t *dg_coo = nullptr;

  // One packed (u << 32 | v), u < v, per undirected edge of the CSR.
  int64_t build(const gsw_host::HostGraph &hg) {
    nvertexs = hg.n;
    std::vector<packed_t> pk;
    pk.reserve((size_t)hg.m / 2 + 1);
    for (int64_t u = 0; u < hg.n; ++u)
      for (int64_t j = hg.row_ptr[u]; j < hg.row_ptr[u + 1]; ++j)
        if (hg.directed || hg.col[j] > u)
          pk.push_back(((packed_t)u << 32) | (uint32_t)hg.col[j]);
    nedges = (int64_t)pk.size();
    H_ERR(hipMalloc((void **)&dg_coo, sizeof(packed_t) * (nedges + 1)));
    if (nedges)
      H_ERR(hipMemcpy(dg_coo, pk.data(), sizeof(packed_t) * nedges,
                      hipMemcpyHostToDevice));
    return (int64_t)sizeof(packed_t) * nedges;
  }
  void release() {
    hipFree(dg_coo);
    dg_coo = nullptr;
  }
  __host__ __device__ __forceinline__ int get_level() const { return level; }
  __host__ __device__ inline void update_level(int inc = 1) { level += inc; }
  __host__ __device__ inline void reset_level() { level = 0; }
};
