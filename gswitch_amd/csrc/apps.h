// Shipped functors and level-loop drivers behind the C-ABI.
//
// Each functor restates the reference application's operator and each
// driver its run_* loop:
//   BFS   application/bfs.cu:16-36, 78-102   (levels, -1 unreached)
//   SSSP  application/sssp.cu:17-53, 106-126 (uint32 distances; the
//         reference's int / INT_MAX become uint32 / UINT32_MAX)
//   CC    application/cc.cu:15-32, 82-108    (edge-centric hooking + pointer
//         jumping; label = smallest vertex id of the component)
//   PR    application/pagerank.cu:18-42, 81-103 (un-normalised form; the
//         accumulator is FP64 and the rank stored per out-degree is FP32;
//         eps = 0 with an iteration cap gives a fixed iteration count)
#pragma once
#include "gswitch/gswitch.h"

namespace gsw_apps {

using GU = device_graph_t<CSR, Empty>;
using GW = device_graph_t<CSR, uint32_t>;
using GC = device_graph_t<COO, Empty>;

struct BFS : Functor<VC, int, Empty, Empty> {
  static constexpr bool level_sync = true;
  __device__ Status filter(int v, GU g) {
    const int lvl = *wa_of(v);
    if (lvl == g.get_level()) return Active;
    if (lvl < 0) return Inactive;
    return Fixed;
  }
  __device__ int emit(int v, Empty *e, GU g) { return g.get_level(); }
  __device__ bool cond(int v, int msg, GU g) { return *wa_of(v) == -1; }
  __device__ bool comp(int *v, int msg, GU g) {
    *v = msg;
    return true;
  }
  __device__ bool compAtomic(int *v, int msg, GU g) {
    *v = msg;
    return true;
  }
};

struct SSSP : Functor<VC, uint32_t, uint32_t, uint32_t> {
  __device__ Status filter(int v, GW g) {
    uint32_t *nd = wa_of(v);
    uint32_t *od = ra_of(v);
    if (*nd == *od) return Inactive;
    if (data.window.in((double)*nd)) {
      *od = *nd;
      return Active;
    }
    return Inactive;
  }
  // saturating: a path longer than 2^32 - 1 counts as no path (UINT32_MAX)
  // instead of wrapping to a short distance
  __device__ uint32_t emit(int v, uint32_t *w, GW g) {
    const uint32_t d = *wa_of(v), s = d + *w;
    return s < d ? UINT32_MAX : s;
  }
  __device__ bool comp(uint32_t *d, uint32_t msg, GW g) {
    if (*d > msg) {
      *d = msg;
      return data.window.in((double)msg);
    }
    return false;
  }
  __device__ bool compAtomic(uint32_t *d, uint32_t msg, GW g) {
    // distances only decrease, so a plain read that is already <= msg
    // proves the atomic would change nothing
    if (*d <= msg) return false;
    const uint32_t old = atomicMin(d, msg);
    return old > msg && data.window.in((double)msg);
  }
  __device__ bool cond(int v, uint32_t msg, GW g) { return *wa_of(v) > msg; }
  __device__ bool exit(int v, GW g) { return *wa_of(v) != *ra_of(v); }
};

struct CC : Functor<EC, int, Empty, Empty> {
  __device__ Status filter(int v, int u, Empty *e) {
    return (*wa_of(v) == *wa_of(u)) ? Fixed : Active;
  }
  __device__ void update(int v, int u, Empty *e) {
    int *rv = wa_of(*wa_of(v));
    int *ru = wa_of(*wa_of(u));
    const int a = *rv, b = *ru;
    if (a < b)
      atomicMin(ru, a);
    else if (a > b)
      atomicMin(rv, b);
  }
};

struct PR : Functor<VC, double, float, Empty> {
  // a message is a stored FP32 rank widened to FP64: it is stored exactly
  // as FP32 where messages are staged per edge (propagation blocking)
  using msg_t = float;
  float damp;
  float eps;
  __device__ Status filter(int v, GU g) {
    double *cache = wa_of(v);
    const double acc = *cache;
    *cache = 0;
    return filter_acc(v, acc, g);
  }
  // the filter given the accumulated value, with WA already at zero: a
  // propagation-blocked gather passes its LDS accumulator and leaves WA alone
  __device__ Status filter_acc(int v, double acc, GU g) {
    float *pr = ra_of(v);
    const int od = g.get_out_degree(v);
    double np = acc * (double)damp + (1.0 - (double)damp);
    if (od) np /= od;
    const float err = (float)((np - (double)*pr) / (double)*pr);
    *pr = (float)np;
    if (err > -eps && err < eps) return Inactive;
    return Active;
  }
  __device__ double emit(int v, Empty *e, GU g) { return (double)*ra_of(v); }
  __device__ bool comp(double *v, double msg, GU g) {
    *v += msg;
    return true;
  }
  __device__ bool compAtomic(double *v, double msg, GU g) {
    atomicAdd(v, msg);
    return true;
  }
};

// ---- Delta-PageRank (reference docs/index.md:175-179, not in its code) ----
// Each vertex keeps its value and the delta it sends this round (already
// divided by its out-degree) in RA; WA accumulates the damped deltas that
// arrive. The filter applies an accumulation once it exceeds eps (the vertex
// becomes active and sends it on next round) and otherwise keeps it
// accumulating, so no mass is lost and the values converge to the PageRank
// fixed point value = (1 - d) + d * sum value[u] / degree[u].
struct dpr_t {
  double pr;
  double send;
};
struct DeltaPR : Functor<VC, double, dpr_t, Empty> {
  float damp;
  double eps;
  __device__ Status filter(int v, GU g) {
    double *acc = wa_of(v);
    dpr_t *s = ra_of(v);
    const double a = *acc;
    if (a > eps || a < -eps) {
      const int od = g.get_out_degree(v);
      s->pr += a;
      s->send = od ? a / (double)od : 0.0;
      *acc = 0;
      return Active;
    }
    s->send = 0;
    return Inactive;
  }
  __device__ double emit(int v, Empty *e, GU g) {
    return (double)damp * ra_of(v)->send;
  }
  __device__ bool comp(double *v, double msg, GU g) {
    *v += msg;
    return true;
  }
  __device__ bool compAtomic(double *v, double msg, GU g) {
    atomicAdd(v, msg);
    return true;
  }
};

// ---- graph colouring (application/gc.cu:15-43) ---------------------------
// Jones-Plassmann rounds with vertex ids as priorities: each round every
// uncoloured vertex pulls the largest id among its uncoloured neighbours
// (WA, reset to -1 by the filter); an uncoloured vertex whose own id is
// larger takes the round number as its colour (RA) and leaves the frontier.
// Colours start at 1; coloured vertices are Fixed.
struct Color : Functor<VC, int, int, Empty> {
  __device__ Status filter(int v, GU g) {
    int *color = ra_of(v);
    int *maxid = wa_of(v);
    if (*color != -1) return Fixed;
    if (*maxid < v) {
      *color = g.get_level();
      return Fixed;
    }
    *maxid = -1;
    return Active;
  }
  __device__ int emit(int v, Empty *e, GU g) { return *ra_of(v) == -1 ? v : -1; }
  __device__ bool cond(int v, int msg, GU g) { return *ra_of(v) == -1; }
  __device__ bool comp(int *v, int msg, GU g) {
    if (*v < msg) *v = msg;
    return true;
  }
  __device__ bool compAtomic(int *v, int msg, GU g) {
    atomicMax(v, msg);
    return true;
  }
};

// ---- betweenness centrality from one root (application/bc.cu:16-94) -----
// Forward pass: level + number of shortest paths (nsp); backward pass:
// Brandes dependency accumulated from the deepest level up. nsp is an FP64
// count (the reference's int overflows on large graphs) and dep is FP64
// (the reference's float).
struct bc_fwd_t {
  int level;  // -1 until reached
  int pad;
  double nsp;
};
struct bc_bwd_t {
  double nsp;
  double dep;
};

struct BCF : Functor<VC, bc_fwd_t, Empty, Empty> {
  static constexpr bool level_sync = true;
  __device__ Status filter(int v, GU g) {
    const int lvl = wa_of(v)->level;
    if (lvl == g.get_level()) return Active;
    if (lvl < 0) return Inactive;
    return Fixed;
  }
  __device__ bc_fwd_t emit(int v, Empty *e, GU g) {
    bc_fwd_t x = *wa_of(v);
    x.level += 1;
    return x;
  }
  __device__ bool cond(int v, bc_fwd_t m, GU g) {
    const int lvl = wa_of(v)->level;
    return lvl < 0 || lvl == g.get_level();
  }
  __device__ bool comp(bc_fwd_t *v, bc_fwd_t m, GU g) {
    v->level = m.level;
    v->nsp += m.nsp;
    return true;
  }
  __device__ bool compAtomic(bc_fwd_t *v, bc_fwd_t m, GU g) {
    v->level = m.level;
    atomicAdd(&v->nsp, m.nsp);
    return true;
  }
};

// ra = forward level. Becoming active finalises dep = nsp * sum over
// children of (1 + dep_c) / nsp_c.
struct BCB : Functor<VC, bc_bwd_t, int, Empty> {
  __device__ Status filter(int v, GU g) {
    const int lvl = *ra_of(v);
    if (lvl == g.get_level()) {
      bc_bwd_t *x = wa_of(v);
      x->dep = x->nsp * x->dep;
      return Active;
    }
    if (lvl < g.get_level()) return Inactive;
    return Fixed;
  }
  __device__ bc_bwd_t emit(int v, Empty *e, GU g) {
    bc_bwd_t x = *wa_of(v);
    x.dep = (x.dep + 1.0) / x.nsp;
    return x;
  }
  __device__ bool cond(int v, bc_bwd_t m, GU g) { return *ra_of(v) == g.get_level(); }
  __device__ bool comp(bc_bwd_t *v, bc_bwd_t m, GU g) {
    v->dep += m.dep;
    return true;
  }
  __device__ bool compAtomic(bc_bwd_t *v, bc_bwd_t m, GU g) {
    atomicAdd(&v->dep, m.dep);
    return true;
  }
};

// ---- per-run engine state ------------------------------------------------
struct engine_t {
  inspector_t inspector;
  selector_t selector;
  executor_t executor;
  feature_t fets;
  config_t conf;
  stat_t stats;
  void release() {
    stats.release();
    executor.release();
  }
};

struct run_opts_t {
  ins_t force;
  int max_levels = 1 << 20;
  bool verbose = false;
  bool sorted = false;
  int twc_small = 32, twc_large = 2048;
  bool mp_pull = true;
  bool pb_pull = false;
  int selector = 0;  // 0 rule trees, 1 learned MLP
};

// Switch the engine to the learned selector (weights on the device and a
// host copy) for a graph with these degree statistics. Called after
// init_conf: the first level's choice (fused or not) is the network's.
inline void enable_learned(engine_t &en, const float *w_dev, const float *w_host,
                           const graph_attr_t &attr, double n, double m,
                           bool fromall, bool toall, bool pr_classes = false) {
  gsw_mlp::sel_dev_t &d = en.stats.mlp;
  d.w = w_dev;
  d.n = (float)n;
  d.m = (float)m;
  d.avg_deg = (float)attr.avg_deg;
  d.std_deg = (float)attr.std_deg;
  d.gini = (float)attr.GI;
  d.mask = pr_classes ? gsw_mlp::PR_CLASSES : gsw_mlp::valid_mask(fromall, toall);
  en.conf.mlp_pr_classes = pr_classes;
  en.stats.mlp_host_w = w_host;
  en.conf.policy_mlp = true;
  en.conf.mlp_class = -1;
  en.conf.conf_fuse_inspect = false;
  en.conf.conf_fusion = false;
}

inline void apply_opts(config_t &conf, const run_opts_t &o) {
  conf.force = o.force;
  conf.conf_sorted = o.sorted;
  conf.twc_small = o.twc_small;
  conf.twc_large = o.twc_large;
  conf.mp_pull = o.mp_pull;
  conf.pb_pull = o.pb_pull;
}

// Generic vertex-centric level loop (application/bfs.cu:87-98 shape).
// `stop_after_inspect(level)` lets PageRank end on a fold pass. `g` keeps
// the level the loop ended on (BC's backward pass starts from it).
template <typename G, typename F, typename Stop>
int run_vc(engine_t &en, active_set_t &as, G &g, F &f, const run_opts_t &o,
           Stop stop_after_inspect) {
  int level = 0;
  for (level = 0; level < o.max_levels; ++level) {
    as.window_moved = false;
    en.inspector.inspect(as, g, f, en.stats, en.fets, en.conf);
    if (as.finish(g, f, en.conf)) break;
    // the inspection drained and the window advanced: inspect again rather
    // than expand an empty frontier (a fused choice for it would let the
    // next finish() absorb the new window's updates unexpanded)
    if (as.take_window_moved()) continue;
    if (stop_after_inspect(level)) break;
    en.selector.select(en.stats, en.fets, en.conf);
    en.executor.filter(as, g, f, en.stats, en.fets, en.conf);
    g.update_level();
    en.executor.expand(as, g, f, en.stats, en.fets, en.conf);
    en.fets.record();  // a training record when a record file is set
    if (o.verbose) {
      const auto &r = en.executor.log.back();
      fprintf(stderr, "level %d frontier %lld ", r.level, r.frontier);
      en.conf.show_hints(std::cerr);
      fprintf(stderr, "\n");
    }
    if (as.finish(g, f, en.conf)) break;
  }
  return level;
}

}  // namespace gsw_apps
