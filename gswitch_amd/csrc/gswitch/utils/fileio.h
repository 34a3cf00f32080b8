// Host graphs of a GSWITCH application.
//
// The reference's host-side types with their names and fields
// (src/utils/fileIO.cuh:69-203, 205-529; src/data_structures/graph.cuh:
// 177-280):
//   edgelist_t<E>          read_mtx(path, directed, with_weight, with_header),
//                          gen_weight(lim), nvertexs / nedges / mean_weight
//   host_graph_t<CSR, E>   nvertexs, nedges, start_pos, odegrees, adj_list,
//                          edge_weights (+ r_* for directed graphs),
//                          build(el, quiet), random_root(), attr
//   host_graph_t<COO, E>   nvertexs, nedges, edge_list, edge_weights
//   graph_t<Fmt, E>        { el, hg, dg }
//   build_graph<C, E>(path, fets, with_header, with_weight, directed)
//   build_graph<C, E>(el, fets)
// Files are parsed by the library's reader (gsw_read_mtx) and CSR arrays
// are built by the library's builder (gsw_host_graph_from_edges), so a file
// gives the same vertex numbering on every path. Offsets are int64:
// start_pos[v] .. start_pos[v + 1] is row v for every v (the reference's
// int32 start_pos has no entry n). A host graph is a cheap handle (copies
// share the arrays), so it can be passed by value as the reference apps do.
// Needs libgswitch_amd.so at link time and include/ on the include path.
#pragma once
#include <cstdlib>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../host/host_graph.h"
#include "../abstraction/features.h"
#include "../data_structures/graph.h"
#include "cmdline.h"
#include "gswitch_amd.h"

namespace gsw_io {

inline void check(gsw_status rc) {
  if (rc != GSW_OK) throw gsw::error(rc, gsw_last_error());
}

// Export a library host graph into a HostGraph (forward CSR only).
inline gsw_host::HostGraph export_graph(gsw_host_graph *h) {
  gsw_host::HostGraph g;
  int64_t n = 0, m = 0;
  int32_t dir = 0, wt = 0;
  check(gsw_host_graph_info(h, &n, &m, &dir, &wt));
  g.n = n;
  g.m = m;
  g.directed = dir != 0;
  g.weighted = wt != 0;
  g.row_ptr.resize((size_t)n + 1);
  g.col.resize((size_t)m);
  if (g.weighted) g.w.resize((size_t)m);
  check(gsw_host_graph_export(h, g.row_ptr.data(), g.col.data(),
                              g.weighted ? g.w.data() : nullptr));
  gsw_host_graph_destroy(h);
  return g;
}

// Reverse CSR (in-edges) of a directed graph, rows sorted.
inline void build_reverse(gsw_host::HostGraph &g) {
  g.r_row_ptr.assign((size_t)g.n + 1, 0);
  for (int32_t u : g.col) ++g.r_row_ptr[(size_t)u + 1];
  for (int64_t v = 0; v < g.n; ++v) g.r_row_ptr[v + 1] += g.r_row_ptr[v];
  std::vector<int64_t> pos(g.r_row_ptr.begin(), g.r_row_ptr.end() - 1);
  g.r_col.resize((size_t)g.m);
  if (g.weighted) g.r_w.resize((size_t)g.m);
  for (int64_t v = 0; v < g.n; ++v)  // ascending sources: rows come out sorted
    for (int64_t j = g.row_ptr[v]; j < g.row_ptr[v + 1]; ++j) {
      const int64_t p = pos[g.col[j]]++;
      g.r_col[p] = (int32_t)v;
      if (g.weighted) g.r_w[p] = g.w[j];
    }
}

}  // namespace gsw_io

template <typename E> struct edgelist_t {
  // edge k is (vvector[2k], vvector[2k + 1]) with weight evector[k]; an
  // undirected graph lists each edge once
  std::vector<int32_t> vvector;
  std::vector<uint32_t> evector;
  int64_t nvertexs = 0;
  int64_t nedges = 0;
  int vmin = 0, vmax = -1;
  bool directed = false;
  bool with_weight = false;
  double mean_weight = 0;

  // Matrix Market / edge-list file (edgelist_t::read_mtx,
  // src/utils/fileIO.cuh:73-146): comment lines skipped, an optional size
  // line, self-loops dropped, ids shifted by the smallest id, duplicates
  // removed. Vertex ids are not renumbered here; host_graph_t::build drops
  // degree-zero vertices where the reference does.
  void read_mtx(const std::string &path, bool directed_ = false,
                bool with_weight_ = false, bool with_header = false) {
    gsw_host_graph *h = nullptr;
    gsw_io::check(gsw_read_mtx(path.c_str(), directed_, with_weight_, with_header,
                               0, &h));
    const gsw_host::HostGraph g = gsw_io::export_graph(h);
    directed = directed_;
    with_weight = g.weighted;
    nvertexs = g.n;
    vmin = 0;
    vmax = (int)g.n - 1;
    vvector.clear();
    evector.clear();
    double tot = 0;
    for (int64_t v = 0; v < g.n; ++v)
      for (int64_t j = g.row_ptr[v]; j < g.row_ptr[v + 1]; ++j) {
        if (!directed && g.col[j] < v) continue;  // each undirected edge once
        vvector.push_back((int32_t)v);
        vvector.push_back(g.col[j]);
        if (with_weight) {
          evector.push_back(g.w[j]);
          tot += g.w[j];
        }
      }
    nedges = (int64_t)vvector.size() / 2;
    mean_weight = (with_weight && nedges) ? tot / (double)nedges : 0;
  }

  // Edge weights in [1, lim] where the file has none (gen_weight,
  // fileIO.cuh:148-166): one seeded draw per edge, the same for both
  // directions of an undirected edge; a weighted list keeps its weights.
  void gen_weight(int lim) {
    if (!with_weight || evector.size() != (size_t)nedges) {
      const uint64_t key = gsw_host::stream_key(0x6A09E667F3BCC908ull,
                                                gsw_host::STREAM_WEIGHT);
      evector.resize((size_t)nedges);
      for (int64_t k = 0; k < nedges; ++k) {
        const uint32_t a = (uint32_t)vvector[2 * k], b = (uint32_t)vvector[2 * k + 1];
        evector[k] = gsw_host::pair_weight(key, a < b ? a : b, a < b ? b : a,
                                           (uint32_t)lim);
      }
      with_weight = true;
    }
    double tot = 0;
    for (uint32_t w : evector) tot += w;
    mean_weight = nedges ? tot / (double)nedges : 0;
  }
};

template <GraphFmt Fmt, typename E> struct host_graph_t {};

template <typename E> struct host_graph_t<CSR, E> {
  static constexpr bool weighted_t = !std::is_same<E, Empty>::value;
  struct store_t {
    gsw_host::HostGraph g;
    std::vector<int32_t> odeg, r_odeg;
    std::vector<E> w, r_w;
  };
  std::shared_ptr<store_t> st;

  int64_t nvertexs = 0;
  int64_t nedges = 0;  // directed CSR entries (both directions counted)
  bool directed = false;
  bool with_weight = false;
  double mean_weight = 0;
  int64_t *start_pos = nullptr;  // [nvertexs + 1]
  int32_t *odegrees = nullptr;
  int32_t *adj_list = nullptr;
  E *edge_weights = nullptr;
  int64_t *r_start_pos = nullptr;  // in-edges of a directed graph
  int32_t *r_odegrees = nullptr;
  int32_t *r_adj_list = nullptr;
  E *r_edge_weights = nullptr;
  graph_attr_t attr;

  // CSR from an edge list (host_graph_t<CSR>::build, fileIO.cuh:210-332):
  // undirected edges stored both ways, duplicates removed, rows sorted. An
  // undirected unweighted graph that is not built quietly also loses its
  // degree-zero vertices, the rest renumbered in order (rinse,
  // fileIO.cuh:399-443).
  void build(const edgelist_t<E> &el, bool quiet = false) {
    const int64_t ne = el.nedges;
    std::vector<int32_t> src((size_t)ne), dst((size_t)ne);
    for (int64_t k = 0; k < ne; ++k) {
      src[k] = el.vvector[2 * k];
      dst[k] = el.vvector[2 * k + 1];
    }
    int64_t n = el.nvertexs;
    if (!weighted_t && !quiet && !el.directed) {
      std::vector<int32_t> id((size_t)n, -1);
      for (int64_t k = 0; k < ne; ++k) id[src[k]] = id[dst[k]] = 0;
      int32_t next = 0;
      for (auto &x : id)
        if (x == 0) x = next++;
      for (int64_t k = 0; k < ne; ++k) {
        src[k] = id[src[k]];
        dst[k] = id[dst[k]];
      }
      n = next;
    }
    const bool use_w = weighted_t && el.with_weight;
    gsw_host_graph *h = nullptr;
    gsw_io::check(gsw_host_graph_from_edges(
        n > 0 ? n : 1, ne, src.data(), dst.data(),
        use_w ? el.evector.data() : nullptr, el.directed, &h));
    adopt(gsw_io::export_graph(h));
  }

  // Take over a host CSR built elsewhere (a generator, the C-ABI).
  void adopt(gsw_host::HostGraph g) {
    if (g.directed && g.r_row_ptr.empty()) gsw_io::build_reverse(g);
    st = std::make_shared<store_t>();
    st->g = std::move(g);
    gsw_host::HostGraph &hg = st->g;
    nvertexs = hg.n;
    nedges = hg.m;
    directed = hg.directed;
    with_weight = hg.weighted;
    auto degrees = [&](const std::vector<int64_t> &rp, std::vector<int32_t> &d) {
      d.resize((size_t)hg.n + 1);
      for (int64_t v = 0; v < hg.n; ++v) d[v] = (int32_t)(rp[v + 1] - rp[v]);
    };
    degrees(hg.row_ptr, st->odeg);
    start_pos = hg.row_ptr.data();
    odegrees = st->odeg.data();
    adj_list = hg.col.data();
    if (directed) {
      degrees(hg.r_row_ptr, st->r_odeg);
      r_start_pos = hg.r_row_ptr.data();
      r_odegrees = st->r_odeg.data();
      r_adj_list = hg.r_col.data();
    } else {
      r_start_pos = start_pos;
      r_odegrees = odegrees;
      r_adj_list = adj_list;
    }
    mean_weight = 0;
    if constexpr (weighted_t) {
      if (with_weight) {
        st->w.assign(hg.w.begin(), hg.w.end());
        double tot = 0;
        for (uint32_t x : hg.w) tot += x;
        mean_weight = hg.m ? tot / (double)hg.m : 0;
        edge_weights = st->w.data();
        if (directed) {
          st->r_w.assign(hg.r_w.begin(), hg.r_w.end());
          r_edge_weights = st->r_w.data();
        } else {
          r_edge_weights = edge_weights;
        }
      }
    }
    attr.fill(hg.row_ptr.data(), hg.n);
  }

  const gsw_host::HostGraph &host() const { return st->g; }

  // A vertex with out-edges, drawn with the C library's random() as the
  // reference does (host_graph_t::random_root, fileIO.cuh:389-397); 0 for a
  // graph without edges.
  int random_root() const {
    if (nedges == 0) return 0;
    for (;;) {
      const int r = (int)(random() % nvertexs);
      if (odegrees[r]) return r;
    }
  }
};

template <typename E> struct host_graph_t<COO, E> {
  // the CSR the packed edge list is cut from
  host_graph_t<CSR, E> csr;
  int64_t nvertexs = 0;
  int64_t nedges = 0;  // undirected edges (each once)
  bool directed = false;
  bool with_weight = false;
  graph_attr_t attr;

  // host_graph_t<COO>::build (fileIO.cuh:471-520): no renumbering.
  void build(const edgelist_t<E> &el, bool quiet = false) {
    csr.build(el, true);
    nvertexs = csr.nvertexs;
    directed = el.directed;
    nedges = directed ? csr.nedges : csr.nedges / 2;
    with_weight = csr.with_weight;
    attr = csr.attr;
    (void)quiet;
  }
  const gsw_host::HostGraph &host() const { return csr.host(); }
};

// graph_t (src/data_structures/graph.cuh:177-181): the edge list, the host
// graph and the device graph of one input.
template <GraphFmt Fmt, typename E> struct graph_t {
  edgelist_t<E> el;
  host_graph_t<Fmt, E> hg;
  device_graph_t<Fmt, E> dg;
  graph_attr_t attr;
  double mean_weight = 0;
  // out-degree of v (the first frontier's workload when v is the root)
  double root_degree(int64_t v) const {
    const auto &h = hg.host();
    return (double)(h.row_ptr[v + 1] - h.row_ptr[v]);
  }
};

// The graph of an edge list, on the device, with the graph-level features
// set (Fake<VC|EC>::func2, src/data_structures/graph.cuh:199-262).
template <Centric C, typename E = Empty>
auto build_graph(const edgelist_t<E> &el, feature_t &fets) {
  constexpr GraphFmt fmt = (C == VC) ? CSR : COO;
  graph_t<fmt, E> g;
  g.el = el;
  g.hg.build(g.el);
  g.attr = g.hg.attr;
  if constexpr (C == VC) {
    g.mean_weight = g.hg.mean_weight;
    g.dg.build(g.hg.host());
    fets.set_graph((double)g.hg.nvertexs,
                   (double)(g.hg.directed ? g.hg.nedges : g.hg.nedges / 2),
                   g.attr, g.mean_weight);
  } else {
    g.dg.build(g.hg.host());
    fets.set_graph((double)g.hg.nvertexs, (double)g.dg.nedges, g.attr);
  }
  return g;
}

// build_graph<C, E>(path, ...) (graph.cuh:268-274): read the file, then as
// above. The weights of a weighted file are kept when E carries them.
template <Centric C, typename E = Empty>
auto build_graph(const std::string &path, feature_t &fets, bool with_header = false,
                 bool with_weight = false, bool directed = false) {
  edgelist_t<E> el;
  el.read_mtx(path, directed, with_weight && !std::is_same<E, Empty>::value,
              with_header);
  return build_graph<C, E>(el, fets);
}

// A graph from a host CSR built elsewhere (generators, the C-ABI).
template <Centric C, typename E = Empty>
auto build_graph(const gsw_host::HostGraph &hg, feature_t &fets) {
  constexpr GraphFmt fmt = (C == VC) ? CSR : COO;
  graph_t<fmt, E> g;
  host_graph_t<CSR, E> csr;
  csr.adopt(hg);
  g.attr = csr.attr;
  if constexpr (C == VC) {
    g.hg = csr;
    g.mean_weight = csr.mean_weight;
    g.dg.build(g.hg.host());
    fets.set_graph((double)hg.n, (double)(hg.directed ? hg.m : hg.m / 2), g.attr,
                   g.mean_weight);
  } else {
    g.hg.csr = csr;
    g.hg.nvertexs = csr.nvertexs;
    g.hg.directed = csr.directed;
    g.hg.nedges = csr.directed ? csr.nedges : csr.nedges / 2;
    g.hg.attr = csr.attr;
    g.dg.build(g.hg.host());
    fets.set_graph((double)hg.n, (double)g.dg.nedges, g.attr);
  }
  return g;
}
