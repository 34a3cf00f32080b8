// C-ABI of the traversal engine (declarations: include/gswitch_amd.h).
#include "../../include/gswitch_amd.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "apps.h"
#include "gswitch/kernels/cc.h"
#include "gswitch/kernels/dobfs.h"
#include "gswitch/kernels/msbfs.h"
#include "host/host_graph.h"

using namespace gsw_apps;

namespace {

thread_local std::string g_last_error;

gsw_status fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

template <typename Fn> gsw_status guarded(Fn fn) {
  try {
    g_last_error.clear();
    fn();
    return GSW_OK;
  } catch (const gsw::error &e) {
    return fail(e.code, e.what());
  } catch (const std::invalid_argument &e) {
    return fail(GSW_E_INVALID, e.what());
  } catch (const std::bad_alloc &) {
    return fail(GSW_E_ALLOC, "host allocation failed");
  } catch (const std::exception &e) {
    return fail(GSW_E_INTERNAL, e.what());
  }
}

void require_gpu() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    throw gsw::error(GSW_E_HIP, "no GPU available");
}

// ---- app-specific init / output kernels ----------------------------------
__global__ void k_init_bfs(int *lvl, int64_t n, int root) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    lvl[v] = (v == root) ? 0 : -1;
}
__global__ void k_set_root_levels(int *lvl, const int32_t *roots, int64_t k) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k;
       i += (int64_t)gridDim.x * blockDim.x)
    lvl[roots[i]] = 0;
}
// End of one captured BFS level that consumed queue `cur`: the level moves
// on only if the level produced a frontier (so *level ends as the number of
// expanded non-empty frontiers), and the consumed size is cleared for the
// level after next.
__global__ void k_level_advance(int *level, int32_t *sizes, int cur) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (sizes[cur ^ 1] != 0) *level += 1;
    sizes[cur] = 0;
  }
}
__global__ void k_init_sssp(uint32_t *wa, uint32_t *ra, int64_t n, int root) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t x = (v == root) ? 0u : UINT32_MAX;
    wa[v] = x;
    ra[v] = x;
  }
}
__global__ void k_init_iota(int *lab, int64_t n) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    lab[v] = (int)v;
}
__global__ void k_init_pr(double *wa, float *ra, const int64_t *row_ptr,
                          int64_t n, float damp) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t od = row_ptr[v + 1] - row_ptr[v];
    wa[v] = 0.0;
    ra[v] = od ? (1.0f - damp) / (float)od : (1.0f - damp);
  }
}
__global__ void k_init_dpr(double *wa, dpr_t *ra, const int64_t *row_ptr,
                           int64_t n, float damp) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t od = row_ptr[v + 1] - row_ptr[v];
    wa[v] = 0.0;
    ra[v].pr = 1.0 - (double)damp;
    ra[v].send = od ? (1.0 - (double)damp) / (double)od : 0.0;
  }
}
__global__ void k_dpr_output(const dpr_t *ra, int64_t n, float *out) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    out[v] = (float)ra[v].pr;
}
__global__ void k_init_bc(bc_fwd_t *wa, int64_t n, int root) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    bc_fwd_t x;
    x.level = (v == root) ? 0 : -1;
    x.pad = 0;
    x.nsp = (v == root) ? 1.0 : 0.0;
    wa[v] = x;
  }
}
// Backward-pass state from the forward result (set_functor,
// application/bc.cu:242-246): ra = level, wa = {nsp, 0}; *max_level =
// deepest level reached.
__global__ void k_bc_backward_init(const bc_fwd_t *fw, bc_bwd_t *bw, int *ra,
                                   int64_t n, int *max_level) {
  int deepest = -1;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const bc_fwd_t x = fw[v];
    ra[v] = x.level;
    bw[v] = bc_bwd_t{x.nsp, 0.0};
    deepest = max(deepest, x.level);
  }
  deepest = gsw::wave_max_any(deepest);
  // read first: most wavefronts hold no new maximum, and atomics from every
  // wavefront onto one address serialise
  if (gsw::lane_id() == 0 && deepest > __hip_atomic_load(max_level, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(max_level, deepest);
}
__global__ void k_bc_output(const bc_fwd_t *fw, const bc_bwd_t *bw, int64_t n,
                            int32_t *level, double *nsp, double *dep) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    level[v] = fw[v].level;
    nsp[v] = bw[v].nsp;
    dep[v] = bw[v].dep;
  }
}
__global__ void k_pr_output(const float *ra, const int64_t *row_ptr,
                            int64_t n, float *out) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t od = row_ptr[v + 1] - row_ptr[v];
    out[v] = od ? (float)((double)ra[v] * (double)od) : ra[v];
  }
}

}  // namespace

struct gsw_host_graph {
  gsw_host::HostGraph g;
};

// Per-application persistent state: frontier + functor arrays, reused by
// every run on the same graph so that runs allocate nothing.
template <typename F> struct app_ctx_t {
  active_set_t as;
  F f{};
  engine_t en;
  bool arch_done = false;
  void release() {
    as.release();
    f.data.release();
    en.release();
  }
};

struct gsw_graph {
  gsw_host::HostGraph hg;
  graph_attr_t attr;
  double mean_weight = 0;
  GU dg;
  GW dgw;
  bool has_w = false;
  GC coo;
  bool has_coo = false;
  std::unique_ptr<app_ctx_t<BFS>> bfs;
  std::unique_ptr<app_ctx_t<SSSP>> sssp;
  std::unique_ptr<app_ctx_t<CC>> cc;
  std::unique_ptr<app_ctx_t<PR>> pr;
  std::unique_ptr<app_ctx_t<Color>> gc;
  std::unique_ptr<app_ctx_t<DeltaPR>> dpr;
  // BC: the forward functor owns the frontier and engine; the backward
  // functor has its own vertex arrays
  std::unique_ptr<app_ctx_t<BCF>> bc;
  BCB bcb{};
  bool bcb_built = false;
  int32_t *d_bc_out = nullptr;  // level | nsp | dep staging (16 B per vertex)
  float *d_out_f = nullptr;
  int32_t *d_flag = nullptr;
  int32_t *h_flag = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<gsw_level_record> records;
  int32_t last_class = -1;  // per-run learned choice (PageRank, CC)
  bool disj_known = false, disj = false;  // batch offsets with no edge between
  uint64_t disj_key = 0;                  //   instances (key: their hash)
  double sssp_window_scale = 1.0;  // SSSP initial window x this
  // 64-source BFS state (allocated on first use)
  uint64_t *ms_buf = nullptr;  // visit | A | B, n each
  gsw::ms_counters_t *ms_cnt = nullptr;
  gsw::ms_counters_t *ms_hcnt = nullptr;
  int32_t *ms_src = nullptr;
  // multi-root BFS state: roots on the device, the device level counter and
  // the captured level loop (instantiated once per launch shape)
  int32_t *d_roots = nullptr;
  int64_t roots_cap = 0;
  int32_t *d_level = nullptr;
  hipGraph_t cap_graph = nullptr;
  hipGraphExec_t cap_exec = nullptr;
  int cap_key[5] = {-1, -1, -1, -1, -1};
  // sampled-hooking CC: labels (parent pointers), the vertex sample for the
  // skipped component, batch instance offsets and the captured kernel
  // sequence of the batched form (instantiated per instance count)
  int32_t *cc_lab = nullptr;
  int32_t *cc_ids = nullptr;   // sampled vertex ids [CC_SAMPLE] + chosen root
  int64_t *inst_off = nullptr;
  int64_t inst_cap = 0;
  hipGraph_t cc_graph = nullptr;
  hipGraphExec_t cc_exec = nullptr;
  int64_t cc_key = -1;
  // device-resident single-query BFS (kernels/dobfs.h): rotating frontier
  // and visited words, heavy lists, control block and its captured loop
  uint64_t *do_bits = nullptr;  // visited [nw] | F [3 nw]
  int32_t *do_heavy = nullptr;
  int64_t *do_hpre = nullptr;
  long long *do_ctrl = nullptr;
  long long *do_hist = nullptr;
  long long *do_hctrl = nullptr;  // pinned copy of the first control words
  hipGraph_t do_graph = nullptr;
  hipGraphExec_t do_exec = nullptr;
  double do_key[5] = {-1, -1, -1, -1, -1};

  void drop_do_capture() {
    if (do_exec) hipGraphExecDestroy(do_exec);
    if (do_graph) hipGraphDestroy(do_graph);
    do_exec = nullptr;
    do_graph = nullptr;
    for (double &x : do_key) x = -1;
  }

  void drop_cc_capture() {
    if (cc_exec) hipGraphExecDestroy(cc_exec);
    if (cc_graph) hipGraphDestroy(cc_graph);
    cc_exec = nullptr;
    cc_graph = nullptr;
    cc_key = -1;
  }

  void drop_capture() {
    if (cap_exec) hipGraphExecDestroy(cap_exec);
    if (cap_graph) hipGraphDestroy(cap_graph);
    cap_exec = nullptr;
    cap_graph = nullptr;
    for (int &x : cap_key) x = -1;
  }

  void release() {
    drop_capture();
    drop_cc_capture();
    drop_do_capture();
    if (do_bits) hipFree(do_bits);
    if (do_heavy) hipFree(do_heavy);
    if (do_hpre) hipFree(do_hpre);
    if (do_ctrl) hipFree(do_ctrl);
    if (do_hist) hipFree(do_hist);
    if (do_hctrl) hipHostFree(do_hctrl);
    if (cc_lab) hipFree(cc_lab);
    if (cc_ids) hipFree(cc_ids);
    if (inst_off) hipFree(inst_off);
    if (d_roots) hipFree(d_roots);
    if (d_level) hipFree(d_level);
    if (bfs) bfs->release();
    if (sssp) sssp->release();
    if (cc) cc->release();
    if (pr) pr->release();
    if (gc) gc->release();
    if (dpr) dpr->release();
    if (bc) bc->release();
    if (bcb_built) bcb.data.release();
    if (d_bc_out) hipFree(d_bc_out);
    if (has_w) hipFree(dgw.dg_edgedata);
    if (has_coo) coo.release();
    dg.release();
    if (d_out_f) hipFree(d_out_f);
    if (d_flag) hipFree(d_flag);
    if (h_flag) hipHostFree(h_flag);
    if (ev0) hipEventDestroy(ev0);
    if (ev1) hipEventDestroy(ev1);
    if (ms_buf) hipFree(ms_buf);
    if (ms_cnt) hipFree(ms_cnt);
    if (ms_hcnt) hipHostFree(ms_hcnt);
    if (ms_src) hipFree(ms_src);
  }
};

namespace {

// Sort every adjacency row ascending (weights move with their columns):
// the pulls that split rows by source range (TWOD slices) and the early-exit
// first probe assume sorted rows. Rows that are already sorted cost one scan.
void sort_rows(gsw_host::HostGraph &hg) {
  std::vector<std::pair<int32_t, uint32_t>> tmp;
  for (int64_t v = 0; v < hg.n; ++v) {
    const int64_t s = hg.row_ptr[v], e = hg.row_ptr[v + 1];
    auto *c = hg.col.data();
    if (std::is_sorted(c + s, c + e)) continue;
    if (hg.w.empty()) {
      std::sort(c + s, c + e);
      continue;
    }
    tmp.clear();
    for (int64_t j = s; j < e; ++j) tmp.emplace_back(c[j], hg.w[j]);
    std::sort(tmp.begin(), tmp.end());
    for (int64_t j = s; j < e; ++j) {
      c[j] = tmp[j - s].first;
      hg.w[j] = tmp[j - s].second;
    }
  }
}

void upload_graph(gsw_graph *G) {
  G->attr.fill(G->hg.row_ptr.data(), G->hg.n);
  G->dg.build(G->hg);
  if (G->hg.weighted) {
    double s = 0;
    for (uint32_t x : G->hg.w) s += x;
    G->mean_weight = G->hg.w.empty() ? 0 : s / (double)G->hg.w.size();
    G->dgw.nvertexs = G->dg.nvertexs;
    G->dgw.nedges = G->dg.nedges;
    G->dgw.directed = G->dg.directed;
    G->dgw.dg_row_ptr = G->dg.dg_row_ptr;
    G->dgw.dg_adj_list = G->dg.dg_adj_list;
    G->dgw.dgr_row_ptr = G->dg.dgr_row_ptr;
    G->dgw.dgr_adj_list = G->dg.dgr_adj_list;
    G->dgw.dgr_first = G->dg.dgr_first;
    G->dgw.dg_odegree = G->dg.dg_odegree;
    G->dgw.dgr_idegree = G->dg.dgr_idegree;
    G->dgw.twod = nullptr;  // TWOD slices are built for unweighted views only
    GSW_REQUIRE(!G->hg.directed, GSW_E_UNSUPPORTED,
                "weighted directed graphs are not supported");
    H_ERR(hipMalloc((void **)&G->dgw.dg_edgedata,
                    sizeof(uint32_t) * (G->hg.m + 1)));
    if (G->hg.m)
      H_ERR(hipMemcpy(G->dgw.dg_edgedata, G->hg.w.data(),
                      sizeof(uint32_t) * G->hg.m, hipMemcpyHostToDevice));
    G->dgw.dgr_edgedata = G->dgw.dg_edgedata;
    G->has_w = true;
  }
  H_ERR(hipMalloc((void **)&G->d_out_f, sizeof(float) * (G->hg.n + 1)));
  H_ERR(hipMalloc((void **)&G->d_flag, sizeof(int32_t) * 4));
  H_ERR(hipHostMalloc((void **)&G->h_flag, sizeof(int32_t) * 4));
  H_ERR(hipEventCreate(&G->ev0));
  H_ERR(hipEventCreate(&G->ev1));
}

run_opts_t to_opts(const gsw_run_opts *o) {
  run_opts_t r;
  if (!o) return r;
  if (o->configs) r.force = parse_configs(o->configs);
  if (o->max_levels > 0) r.max_levels = o->max_levels;
  r.verbose = o->verbose != 0;
  if (o->twc_small > 0) r.twc_small = o->twc_small;
  if (o->twc_large > 0) r.twc_large = o->twc_large;
  GSW_REQUIRE(o->pull_kernel >= 0 && o->pull_kernel <= 2, GSW_E_INVALID,
              "pull_kernel must be 0, 1 or 2");
  r.mp_pull = o->pull_kernel != 1;
  r.pb_pull = o->pull_kernel == 2;
  r.selector = o->selector;
  return r;
}

// options of traversals whose variant choice stays with the rule trees (the
// learned selectors are trained on BFS, SSSP and PageRank runs)
run_opts_t to_opts_rule(const gsw_run_opts *o) {
  run_opts_t r = to_opts(o);
  GSW_REQUIRE(r.selector == 0, GSW_E_UNSUPPORTED,
              "the learned selector is available for BFS, SSSP and PageRank only");
  return r;
}

template <typename F> void prepare_engine(app_ctx_t<F> &c, const run_opts_t &o,
                                          bool profile) {
  c.en.conf = config_t{};
  const feature_t keep = c.en.fets;
  c.en.fets = feature_t{};
  c.en.fets.cap = keep.cap;
  c.en.fets.register_lim = keep.register_lim;
  c.en.fets.thread_lim = keep.thread_lim;
  c.en.fets.sm_num = keep.sm_num;
  c.en.stats.reset();
  c.en.stats.build();
  c.en.inspector = inspector_t{};
  c.en.executor.clear_log();
  c.en.executor.profile = profile;
  apply_opts(c.en.conf, o);
}

template <typename F> void arch_once(app_ctx_t<F> &c) {
  if (!c.arch_done) {
    c.en.fets.architecture_features();
    c.arch_done = true;
  }
}

void collect_records(gsw_graph *G, executor_t &ex, active_set_t &as) {
  ex.finalize_log(as);
  G->records.clear();
  for (const auto &r : ex.log) {
    gsw_level_record o{};
    o.level = r.level;
    o.dir = (int8_t)r.dir;
    o.fmt = (int8_t)r.fmt;
    o.lb = (int8_t)r.lb;
    o.fused = (int8_t)r.fused;
    o.sorted = (int8_t)r.sorted;
    o.kernel = (int8_t)r.kernel;
    o.frontier = r.frontier;
    o.active = r.active;
    o.inactive = r.inactive;
    o.edges = r.edges;
    o.verts = r.verts;
    o.ms = r.ms;
    G->records.push_back(o);
  }
}

template <typename Gv> struct view_t {
  const gsw_host::HostGraph *hg;
  graph_attr_t attr;
  double mean_weight;
  double root_degree(int64_t v) const {
    return (double)(hg->row_ptr[v + 1] - hg->row_ptr[v]);
  }
};

double elapsed(gsw_graph *G) {
  float ms = 0;
  H_ERR(hipEventSynchronize(G->ev1));
  H_ERR(hipEventElapsedTime(&ms, G->ev0, G->ev1));
  return ms;
}

// ---- learned selector state (process-wide, one device) ---------------------
// one network per application (GSW_APP_*), each trained on that
// application's own per-level timings
float *g_mlp_dev[GSW_APPS] = {};
std::vector<float> g_mlp_host[GSW_APPS];
// class trees per application (selector 2)
gsw_tree::table_t g_tree[GSW_APPS];

template <typename F>
void maybe_learned(app_ctx_t<F> &c, const run_opts_t &o, const gsw_graph *G,
                   bool fromall, bool toall, int app) {
  const bool pr = app == GSW_APP_PR;
  if (o.selector == 0) return;
  GSW_REQUIRE(o.selector == 1 || o.selector == 2, GSW_E_INVALID, "unknown selector");
  if (o.selector == 2) {
    GSW_REQUIRE(!g_tree[app].empty(), GSW_E_INVALID,
                "class-tree selector requested but no tree is loaded for this "
                "application");
    enable_learned(c.en, nullptr, nullptr, G->attr, (double)G->hg.n,
                   (double)G->hg.m, fromall, toall, pr);
    c.en.stats.tree = &g_tree[app];
    return;
  }
  GSW_REQUIRE(g_mlp_dev[app], GSW_E_INVALID,
              "learned selector requested but no weights are loaded for this "
              "application");
  enable_learned(c.en, g_mlp_dev[app], g_mlp_host[app].data(), G->attr,
                 (double)G->hg.n, (double)G->hg.m, fromall, toall, pr);
}

// Host-side choice for an application that chooses once per run from the
// inputs of an everything-active first level (n, m, degree statistics,
// level 0, all n vertices active, push workload m), as the engine's first
// round of a from-all run computes them. -1: a class tree's leaf outside
// `mask` (the caller keeps its default).
int choose_once(const gsw_graph *G, int app, int selector, uint32_t mask) {
  GSW_REQUIRE(selector == 1 || selector == 2, GSW_E_INVALID, "unknown selector");
  gsw_mlp::counts_t c;
  c.n = (float)G->hg.n;
  c.m = (float)G->hg.m;
  c.avg_deg = (float)G->attr.avg_deg;
  c.std_deg = (float)G->attr.std_deg;
  c.gini = (float)G->attr.GI;
  c.level = 0.f;
  c.active = (float)G->hg.n;
  c.inactive = 0.f;
  c.push_wl = (float)G->hg.m;
  c.pull_wl = 0.f;
  c.max_deg_active = (float)G->hg.m;
  float x[gsw_mlp::NIN], y[gsw_mlp::NOUT];
  gsw_mlp::features(c, x);
  if (selector == 2) {
    GSW_REQUIRE(!g_tree[app].empty(), GSW_E_INVALID,
                "class-tree selector requested but no tree is loaded for this "
                "application");
    const int k = g_tree[app].eval(x);
    return (k >= 0 && k < gsw_mlp::NOUT && ((mask >> k) & 1u)) ? k : -1;
  }
  GSW_REQUIRE(g_mlp_dev[app], GSW_E_INVALID,
              "learned selector requested but no weights are loaded for this "
              "application");
  gsw_mlp::forward_host(g_mlp_host[app].data(), x, y);
  return gsw_mlp::argmin_valid(y, mask);
}

app_ctx_t<BFS> &bfs_ctx(gsw_graph *G) {
  if (!G->bfs) {
    G->bfs = std::make_unique<app_ctx_t<BFS>>();
    G->bfs->as.build(G->hg.n, Normal);
    G->bfs->f.data.build(G->hg.n);
  }
  return *G->bfs;
}

// Validate, deduplicate and upload the roots of a multi-root run; returns
// the distinct roots (ascending) and their total out-degree.
std::vector<int32_t> upload_roots(gsw_graph *G, const int32_t *roots,
                                  int32_t nroots, double &workload) {
  GSW_REQUIRE(roots && nroots >= 1, GSW_E_INVALID, "no roots");
  std::vector<int32_t> r(roots, roots + nroots);
  for (int32_t v : r)
    GSW_REQUIRE(v >= 0 && v < G->hg.n, GSW_E_INVALID, "root out of range");
  std::sort(r.begin(), r.end());
  r.erase(std::unique(r.begin(), r.end()), r.end());
  if ((int64_t)r.size() > G->roots_cap) {
    if (G->d_roots) H_ERR(hipFree(G->d_roots));
    G->d_roots = nullptr;
    H_ERR(hipMalloc((void **)&G->d_roots, sizeof(int32_t) * r.size()));
    G->roots_cap = (int64_t)r.size();
  }
  H_ERR(hipMemcpy(G->d_roots, r.data(), sizeof(int32_t) * r.size(),
                  hipMemcpyHostToDevice));
  workload = 0;
  for (int32_t v : r)
    workload += (double)(G->hg.row_ptr[v + 1] - G->hg.row_ptr[v]);
  return r;
}

// BFS levels: -1 everywhere, 0 on the roots (one root, or k device roots)
void init_bfs_levels(gsw_graph *G, int *lvl, int root, const int32_t *d_roots,
                     int64_t k) {
  hipStream_t st = gsw::stream();
  hipLaunchKernelGGL(k_init_bfs, dim3(gsw::grid_for(G->hg.n)), dim3(gsw::BLOCK),
                     0, st, lvl, G->hg.n, d_roots ? -1 : root);
  if (d_roots)
    hipLaunchKernelGGL(k_set_root_levels, dim3(gsw::grid_for(k)),
                       dim3(gsw::BLOCK), 0, st, lvl, d_roots, k);
  H_ERR(hipGetLastError());
}

// The per-level engine loop for BFS from one root (d_roots == nullptr) or
// from k distinct device roots.
int bfs_engine_run(gsw_graph *G, const gsw_run_opts *opts, int32_t root,
                   const int32_t *d_roots, int64_t k, double workload) {
  const run_opts_t o = to_opts(opts);
  auto &c = bfs_ctx(G);
  arch_once(c);
  prepare_engine(c, o, !opts || opts->profile);
  feature_t &fets = c.en.fets;
  fets.centric = VC;
  fets.pattern = Idem;
  fets.fromall = false;
  fets.toall = false;
  fets.use_root = root;
  if (d_roots) {
    fets.first_active = (double)k;
    fets.first_workload = workload;
  }
  fets.set_graph((double)G->hg.n, (double)(G->hg.m / 2), G->attr);
  c.en.conf.conf_pruning = true;
  view_t<GU> gv{&G->hg, G->attr, 0.0};
  init_conf(c.en.stats, fets, c.en.conf, gv, c.f);
  maybe_learned(c, o, G, false, false, GSW_APP_BFS);

  H_ERR(hipEventRecord(G->ev0, gsw::stream()));
  c.as.reset();
  if (d_roots)
    c.as.init_roots(d_roots, k, true);
  else
    c.as.init(root, true);
  init_bfs_levels(G, c.f.data.dg_wa, root, d_roots, k);
  GU g = G->dg;
  const int levels = run_vc(c.en, c.as, g, c.f, o, [](int) { return false; });
  H_ERR(hipEventRecord(G->ev1, gsw::stream()));
  return levels;
}

// Multi-root BFS as a captured level loop: every level is a fused push
// expand over the unsorted queue (WM, or TWC degree bins) followed by
// k_level_advance; the BFS level lives in device memory, so `per_launch`
// levels replay as one hipGraph with no host round trip between them. The
// host reads the next queue length once per replay and stops at zero.
int bfs_captured_run(gsw_graph *G, int64_t k, LB lb, int small, int large,
                     int grid, int per_launch, int max_levels) {
  GSW_REQUIRE(lb == WM || lb == TWC, GSW_E_UNSUPPORTED,
              "captured BFS supports the WM and TWC expands");
  GSW_REQUIRE(per_launch >= 2 && per_launch % 2 == 0 && grid >= 1 &&
                  grid <= gsw::MAX_GRID && small >= 1 && large > small,
              GSW_E_INVALID, "invalid captured-loop shape");
  // k_push_wm scans a wavefront's 64 degrees in int32
  GSW_REQUIRE(lb != WM || G->attr.max_deg * gsw::WAVE <= (double)INT32_MAX,
              GSW_E_UNSUPPORTED,
              "degrees too large for the captured WM expand; use TWC");
  auto &c = bfs_ctx(G);
  active_set_t &as = c.as;
  GSW_REQUIRE(as.queue.mode == Normal && as.queue.buf[1], GSW_E_INTERNAL,
              "captured BFS needs double-buffered queues");
  GSW_REQUIRE(k <= as.queue.capacity, GSW_E_OVERFLOW, "too many roots");
  if (lb == TWC) as.bins.ensure(as.size);
  if (!G->d_level) H_ERR(hipMalloc((void **)&G->d_level, sizeof(int32_t) * 2));
  hipStream_t st = gsw::stream();

  H_ERR(hipEventRecord(G->ev0, st));
  as.reset();
  as.set_fmt(Queue);
  H_ERR(hipMemcpyAsync(as.queue.buf[0], G->d_roots, sizeof(int32_t) * k,
                       hipMemcpyDeviceToDevice, st));
  gsw::set_one_async<int32_t>(as.queue.sizes, (int32_t)k);
  gsw::set_one_async<int32_t>(G->d_level, 1);
  init_bfs_levels(G, c.f.data.dg_wa, -1, G->d_roots, k);

  const int key[5] = {(int)lb, small, large, grid, per_launch};
  if (!G->cap_exec || std::memcmp(key, G->cap_key, sizeof(key)) != 0) {
    G->drop_capture();
    config_t conf{};
    conf.conf_dir = Push;
    conf.conf_lb = lb;
    conf.conf_fuse_inspect = true;
    conf.conf_pruning = true;
    conf.conf_toall = false;
    conf.conf_target = Active;
    GU g = G->dg;
    g.level_ptr = G->d_level;
    using Bd = gsw::push_body<GU, BFS, true>;
    H_ERR(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      for (int j = 0; j < per_launch; ++j) {
        active_set_t a = as;
        a.queue.cursor = j & 1;
        Bd body{a, g, c.f, conf};
        if (lb == TWC) {
          gsw::fill_async<int32_t>(as.bins.sizes, 0, 4);
          hipLaunchKernelGGL(gsw::k_twc_bin<GU>, dim3(grid), dim3(gsw::BLOCK), 0,
                             st, a, g, Active, false, small, large);
          hipLaunchKernelGGL(gsw::k_push_twc<Bd>, dim3(grid), dim3(gsw::BLOCK),
                             0, st, body);
        } else {
          hipLaunchKernelGGL(gsw::k_push_wm<Bd>, dim3(grid), dim3(gsw::BLOCK), 0,
                             st, body, gsw::WAVE);
        }
        hipLaunchKernelGGL(k_level_advance, dim3(1), dim3(64), 0, st,
                           G->d_level, as.queue.sizes, j & 1);
        H_ERR(hipGetLastError());
      }
    } catch (...) {
      hipGraph_t dead = nullptr;
      (void)hipStreamEndCapture(st, &dead);
      if (dead) (void)hipGraphDestroy(dead);
      throw;
    }
    H_ERR(hipStreamEndCapture(st, &G->cap_graph));
    H_ERR(hipGraphInstantiate(&G->cap_exec, G->cap_graph, nullptr, nullptr, 0));
    std::memcpy(G->cap_key, key, sizeof(key));
  }

  int32_t *h = G->h_flag;
  for (int done = 0; done < max_levels; done += per_launch) {
    H_ERR(hipGraphLaunch(G->cap_exec, st));
    H_ERR(hipMemcpyAsync(h, as.queue.sizes, sizeof(int32_t),
                         hipMemcpyDeviceToHost, st));
    H_ERR(hipMemcpyAsync(h + 1, as.dg_ctrl + CTRL_OVERFLOW, sizeof(int32_t),
                         hipMemcpyDeviceToHost, st));
    gsw::sync();
    GSW_REQUIRE(h[1] == 0, GSW_E_OVERFLOW, "frontier queue overflow");
    if (h[0] == 0) break;
  }
  H_ERR(hipEventRecord(G->ev1, st));
  H_ERR(hipMemcpy(h, G->d_level, sizeof(int32_t), hipMemcpyDeviceToHost));
  return h[0];
}

int default_capture_grid(const gsw_graph *G) {
  return gsw::grid_for(G->hg.n, gsw::BLOCK, gsw::MAX_GRID);
}

ins_t ins_of_class(int cls) {
  const gsw_mlp::variant_t v = gsw_mlp::variant_of_class(cls);
  ins_t f;
  f.has_direction = f.has_asfmt = f.has_lb = f.has_fusion = f.has_sorted = true;
  f.d = v.dir;
  f.asf = v.fmt;
  f.lb = v.lb;
  f.fusion = v.fused ? Fused : Standalone;
  f.sorted = v.sorted;
  return f;
}

// Fresh per-level engine state for a BFS level probe (forced class `cls`,
// or the rule selector when cls < 0): the level is inspected as a
// standalone level whatever came before it.
app_ctx_t<SSSP> &sssp_ctx(gsw_graph *G) {
  if (!G->sssp) {
    G->sssp = std::make_unique<app_ctx_t<SSSP>>();
    G->sssp->as.build(G->hg.n, Normal);
    G->sssp->f.data.build(G->hg.n);
  }
  return *G->sssp;
}

// The SSSP engine before its first level (run_sssp, application/sssp.cu:
// 106-126): Mono pattern, to-all, window sized by init_conf.
void sssp_prepare(gsw_graph *G, app_ctx_t<SSSP> &c, const run_opts_t &o,
                  int root, bool profile) {
  prepare_engine(c, o, profile);
  feature_t &fets = c.en.fets;
  fets.centric = VC;
  fets.pattern = Mono;
  fets.fromall = false;
  fets.toall = true;
  fets.use_root = root;
  fets.set_graph((double)G->hg.n, (double)(G->hg.m / 2), G->attr,
                 G->mean_weight);
  c.en.conf.conf_window = true;
  c.f.data.window = window_t{};
  view_t<GW> gv{&G->hg, G->attr, G->mean_weight};
  init_conf(c.en.stats, fets, c.en.conf, gv, c.f);
  // the graph's window scale (gsw_sssp_set_window / gsw_autotune_sssp)
  // multiplies init_conf's initial window; distances do not depend on it
  if (G->sssp_window_scale != 1.0)
    c.f.data.window.set_init_winsize(c.f.data.window.wsz_lim * G->sssp_window_scale);
}

// The engine for SSSP level `lvl` in the state the level loop carries into
// it (features of the previous level, the window), with class `cls` forced
// or the rule selector (cls < 0). The level is inspected as a standalone
// level whatever came before it; level 0 is the first round (the root).
void sssp_level_setup(gsw_graph *G, app_ctx_t<SSSP> &c, int root, int cls,
                      int lvl, const feature_t &fets, const window_t &win) {
  run_opts_t o;
  if (cls >= 0) o.force = ins_of_class(cls);
  sssp_prepare(G, c, o, root, false);
  c.en.fets = fets;
  c.f.data.window = win;
  c.en.conf.conf_first_round = (lvl == 0);
  c.en.conf.conf_fuse_inspect = false;
  c.en.conf.conf_fusion = false;
  c.as.reset();
  if (lvl == 0) c.as.init(root, true);
}

// Cost of SSSP level `lvl` with class `cls` from the current distance
// arrays, attributed as bfs_probe_level attributes BFS levels.
double sssp_probe_level(gsw_graph *G, app_ctx_t<SSSP> &c, int root, int cls,
                        int lvl, const feature_t &fets, const window_t &win) {
  sssp_level_setup(G, c, root, cls, lvl, fets, win);
  engine_t &en = c.en;
  GW g = G->dgw;
  g.level = lvl;
  const bool fused = gsw_mlp::variant_of_class(cls).fused;
  en.inspector.inspect(c.as, g, c.f, en.stats, en.fets, en.conf);
  if (c.as.finish(g, c.f, en.conf)) return 0.0;
  gsw::sync();
  double t0 = gsw_now_ms();
  en.selector.select(en.stats, en.fets, en.conf);
  en.executor.filter(c.as, g, c.f, en.stats, en.fets, en.conf);
  if (fused) {
    gsw::sync();
    t0 = gsw_now_ms();
  }
  g.update_level();
  en.executor.profile = true;
  en.executor.expand(c.as, g, c.f, en.stats, en.fets, en.conf);
  if (!fused) {
    en.inspector.inspect(c.as, g, c.f, en.stats, en.fets, en.conf);
  } else {
    const double nq = (double)c.as.known_active;
    const double fr = (double)en.executor.log.back().frontier;
    if (nq > 0 && nq * gsw_mlp::INSPECT_FRACTION >= (double)c.as.size &&
        nq >= gsw_mlp::INSPECT_GROWTH * fr) {
      en.conf.conf_fuse_inspect = false;
      en.inspector.inspect(c.as, g, c.f, en.stats, en.fets, en.conf);
    }
  }
  gsw::sync();
  return gsw_now_ms() - t0;
}

void bfs_level_setup(gsw_graph *G, app_ctx_t<BFS> &c, int root, int cls) {
  run_opts_t o;
  if (cls >= 0) o.force = ins_of_class(cls);
  prepare_engine(c, o, false);
  feature_t &fets = c.en.fets;
  fets.centric = VC;
  fets.pattern = Idem;
  fets.fromall = false;
  fets.toall = false;
  fets.use_root = root;
  fets.set_graph((double)G->hg.n, (double)(G->hg.m / 2), G->attr);
  c.en.conf.conf_pruning = true;
  c.en.conf.conf_first_round = false;
  c.en.conf.conf_fuse_inspect = false;
  c.as.reset();
}

// Cost of running BFS level `lvl` (from the current levels array) with
// class `cls`, in wall-clock ms, attributed the way the level loop incurs
// it: a standalone level pays its filter, its expand and the inspection of
// the next level (nothing extra when the expand already inspected it, as a
// level-synchronous TM pull does); a fused level pays its expand and the
// read of the next queue length. This level's own inspection belongs to the
// level before it and is not timed. A fused level whose next frontier the
// learned engine's gate hands to the inspector (large and grown, the
// INSPECT_FRACTION / INSPECT_GROWTH test) also pays that inspection.
double bfs_probe_level(gsw_graph *G, app_ctx_t<BFS> &c, int root, int cls,
                       int lvl) {
  bfs_level_setup(G, c, root, cls);
  engine_t &en = c.en;
  GU g = G->dg;
  g.level = lvl;
  const bool fused = gsw_mlp::variant_of_class(cls).fused;
  en.inspector.inspect(c.as, g, c.f, en.stats, en.fets, en.conf);
  if (c.as.finish(g, c.f, en.conf)) return 0.0;
  gsw::sync();
  double t0 = gsw_now_ms();
  en.selector.select(en.stats, en.fets, en.conf);
  en.executor.filter(c.as, g, c.f, en.stats, en.fets, en.conf);
  if (fused) {
    gsw::sync();
    t0 = gsw_now_ms();
  }
  g.update_level();
  en.executor.profile = true;
  en.executor.expand(c.as, g, c.f, en.stats, en.fets, en.conf);
  if (!fused) {
    en.inspector.inspect(c.as, g, c.f, en.stats, en.fets, en.conf);
  } else {
    // a next frontier that the learned engine's gate hands to the inspector
    // costs that inspection too (the level after a fused choice pays it)
    const double nq = (double)c.as.known_active;
    const double fr = (double)en.executor.log.back().frontier;
    if (nq > 0 && nq * gsw_mlp::INSPECT_FRACTION >= (double)c.as.size &&
        nq >= gsw_mlp::INSPECT_GROWTH * fr) {
      en.conf.conf_fuse_inspect = false;
      en.inspector.inspect(c.as, g, c.f, en.stats, en.fets, en.conf);
    }
  }
  gsw::sync();
  const double wall = gsw_now_ms() - t0;
  static const bool verbose = std::getenv("GSW_EXPLORE_VERBOSE") != nullptr;
  if (verbose) {
    en.executor.finalize_log(c.as);
    const auto &r = en.executor.log.back();
    fprintf(stderr, "probe level %d class %d: dir %d fmt %d lb %d fused %d "
                    "frontier %lld edges %lld expand %.3f ms, level %.3f ms\n",
            lvl, cls, (int)r.dir, (int)r.fmt, (int)r.lb, (int)r.fused,
            (long long)r.frontier, (long long)r.edges, r.ms, wall);
  }
  return wall;
}

// ---- device-resident single-query BFS (kernels/dobfs.h) --------------------
double env_num(const char *name, double dflt) {
  const char *e = std::getenv(name);
  if (!e || !*e) return dflt;
  char *end = nullptr;
  const double x = std::strtod(e, &end);
  return (end && *end == 0) ? x : dflt;
}

// The loop's view of graph G (its buffers are allocated on first use and
// kept for the graph's lifetime, so a captured loop stays valid).
gsw::dobfs_t do_view(gsw_graph *G) {
  using namespace gsw;
  const int64_t n = G->hg.n, nw = (n + 63) / 64;
  if (!G->do_bits) {
    H_ERR(hipMalloc((void **)&G->do_bits, sizeof(uint64_t) * 4 * (nw + 1)));
    H_ERR(hipMalloc((void **)&G->do_heavy, sizeof(int32_t) * 2 * DO_HCAP));
    H_ERR(hipMalloc((void **)&G->do_hpre, sizeof(int64_t) * (DO_HCAP + 1)));
    H_ERR(hipMalloc((void **)&G->do_ctrl, sizeof(long long) * DO_TOTAL));
    H_ERR(hipMalloc((void **)&G->do_hist, sizeof(long long) * DO_HIST * DO_HIST_W));
    H_ERR(hipHostMalloc((void **)&G->do_hctrl, sizeof(long long) * DO_CTRL_WORDS));
  }
  dobfs_t d;
  d.n = n;
  d.nw = nw;
  d.m = G->hg.m;
  d.rows = G->dg.dg_row_ptr;
  d.cols = G->dg.dg_adj_list;
  d.rrows = G->dg.dgr_row_ptr;
  d.rcols = G->dg.dgr_adj_list;
  d.first = G->dg.dgr_first;
  d.odeg = G->dg.dg_odegree;
  d.ideg = G->dg.dgr_idegree;
  d.lv = bfs_ctx(G).f.data.dg_wa;
  d.vis = G->do_bits;
  d.F = G->do_bits + nw;
  d.heavy = G->do_heavy;
  d.hpre = G->do_hpre;
  d.ctrl = G->do_ctrl;
  d.hist = G->do_hist;
  // switching thresholds and the heavy-vertex degree (defaults measured on
  // MI355X; the variables exist for sweeps)
  d.alpha = (float)env_num("GSW_DO_ALPHA", 14.0);
  d.beta = (float)env_num("GSW_DO_BETA", 24.0);
  d.heavy_deg = (int)env_num("GSW_DO_HEAVY", 4096.0);
  // a wavefront scans 64 degrees of at most heavy_deg each in int32
  GSW_REQUIRE(d.alpha > 0 && d.beta > 0 && d.heavy_deg >= 64 &&
                  d.heavy_deg <= INT32_MAX / WAVE,
              GSW_E_INVALID, "invalid direction-switch parameters");
  return d;
}

// One BFS from `root` with the direction chosen on the device: a reset
// launch, then `per_launch` captured levels (expand + decide) per hipGraph
// replay and one read of the control block per replay. Returns the number
// of non-empty frontiers expanded.
// root < 0: the `nroots` distinct roots at G->d_roots (upload_roots), every
// vertex's level its distance to the nearest root.
int bfs_device_run(gsw_graph *G, int32_t root, int per_launch, int grid, int max_levels,
                   int64_t nroots = 0) {
  using namespace gsw;
  GSW_REQUIRE(per_launch >= 1 && per_launch <= 256 && grid >= 1 && grid <= MAX_GRID,
              GSW_E_INVALID, "invalid device-loop shape");
  GSW_REQUIRE(root >= 0 || (nroots >= 1 && G->d_roots), E_INTERNAL_CHECK, "no roots");
  const dobfs_t d = do_view(G);
  hipStream_t st = stream();
  H_ERR(hipEventRecord(G->ev0, st));
  hipLaunchKernelGGL(k_do_reset, dim3(grid_for(MAX(d.n, d.nw))), dim3(BLOCK), 0, st, d,
                     root);
  if (root < 0)
    hipLaunchKernelGGL(k_do_seed, dim3(grid_for(nroots)), dim3(BLOCK), 0, st, d,
                       (const int32_t *)G->d_roots, nroots);
  H_ERR(hipGetLastError());
  const double key[5] = {(double)per_launch, (double)grid, (double)d.heavy_deg,
                         (double)d.alpha, (double)d.beta};
  if (!G->do_exec || std::memcmp(key, G->do_key, sizeof(key)) != 0) {
    G->drop_do_capture();
    H_ERR(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      for (int j = 0; j < per_launch; ++j) {
        hipLaunchKernelGGL(k_do_expand, dim3(grid), dim3(BLOCK), 0, st, d);
        hipLaunchKernelGGL(k_do_decide, dim3(1), dim3(DO_DECIDE_BLOCK), 0, st, d);
        H_ERR(hipGetLastError());
      }
    } catch (...) {
      hipGraph_t dead = nullptr;
      (void)hipStreamEndCapture(st, &dead);
      if (dead) (void)hipGraphDestroy(dead);
      throw;
    }
    H_ERR(hipStreamEndCapture(st, &G->do_graph));
    H_ERR(hipGraphInstantiate(&G->do_exec, G->do_graph, nullptr, nullptr, 0));
    std::memcpy(G->do_key, key, sizeof(key));
  }
  const long long *h = G->do_hctrl;
  bool halted = false;
  for (int done = 0; done < max_levels && !halted; done += per_launch) {
    H_ERR(hipGraphLaunch(G->do_exec, st));
    H_ERR(hipMemcpyAsync(G->do_hctrl, G->do_ctrl, sizeof(long long) * DO_CTRL_WORDS,
                         hipMemcpyDeviceToHost, st));
    sync();
    halted = h[DO_HALT] != 0;
  }
  H_ERR(hipEventRecord(G->ev1, st));
  const int levels = halted ? (int)h[DO_LEVELS] : (int)h[DO_LEVEL];
  // one record per expanded level from the device history: top-down levels
  // are wavefront-mapped pushes over the frontier bitmap, bottom-up levels
  // thread-mapped pulls, both with the inspection folded in
  G->records.clear();
  const int nrec = MIN(levels, DO_HIST);
  if (nrec > 0) {
    std::vector<long long> hv((size_t)nrec * DO_HIST_W);
    H_ERR(hipMemcpy(hv.data(), G->do_hist, sizeof(long long) * hv.size(),
                    hipMemcpyDeviceToHost));
    for (int L = 0; L < nrec; ++L) {
      const long long *r = &hv[(size_t)L * DO_HIST_W];
      gsw_level_record rec{};
      rec.level = L;
      rec.dir = r[0] ? 1 : 0;
      rec.fmt = 1;
      rec.lb = r[0] ? 0 : 1;
      rec.fused = 1;
      rec.frontier = r[1];
      rec.active = r[1];
      rec.edges = r[2];
      rec.verts = r[3];
      G->records.push_back(rec);
    }
  }
  return levels;
}

// default workgroups of the device loop: one frontier word per wavefront
// up to the full grid
int default_do_grid(const gsw_graph *G) {
  const int64_t nw = (G->hg.n + 63) / 64;
  return gsw::grid_for(nw * gsw::WAVE);
}

// ---- connected components by sampled hooking (kernels/cc.h) --------------
constexpr int CC_SAMPLE = 1024;

void cc_buffers(gsw_graph *G) {
  if (G->cc_lab) return;
  H_ERR(hipMalloc((void **)&G->cc_lab, sizeof(int32_t) * (G->hg.n + 1)));
  H_ERR(hipMalloc((void **)&G->cc_ids, sizeof(int32_t) * (CC_SAMPLE + 64)));
}

// init, two sampling rounds, compress (phases 1-2) into lab[n]
void cc_sample_phase(const gsw_graph *G, int32_t *lab, hipStream_t st) {
  using namespace gsw;
  const int64_t n = G->hg.n;
  const int grid = grid_for(n);
  hipLaunchKernelGGL(k_cc_init, dim3(grid), dim3(BLOCK), 0, st, lab, n);
  // a compress after each round: sampled links on a grid build chains
  // thousands long, which the second round's links would walk (4096-square
  // grid: 107 ms with one compress at the end, 3.3 ms with two)
  for (int r = 0; r < 2; ++r) {
    hipLaunchKernelGGL(k_cc_sample, dim3(grid), dim3(BLOCK), 0, st,
                       (const int64_t *)G->dg.dg_row_ptr,
                       (const int32_t *)G->dg.dg_adj_list, n, r, lab,
                       (const int *)nullptr);
    hipLaunchKernelGGL(k_cc_compress, dim3(grid), dim3(BLOCK), 0, st, lab, n);
  }
  H_ERR(hipGetLastError());
}

// the remaining edges outside the skipped set, then compress (phases 3-4)
void cc_finish_phase(const gsw_graph *G, int32_t *lab, int skip_root,
                     const int *skip_dev, const int64_t *off, int64_t count,
                     hipStream_t st) {
  using namespace gsw;
  const int64_t n = G->hg.n;
  hipLaunchKernelGGL(k_cc_rest, dim3(grid_for((n + 63) / 64 * 64)), dim3(BLOCK), 0,
                     st, (const int64_t *)G->dg.dg_row_ptr,
                     (const int32_t *)G->dg.dg_adj_list, n, 2, skip_root, skip_dev,
                     off, count, lab, 0);
  hipLaunchKernelGGL(k_cc_compress, dim3(grid_for(n)), dim3(BLOCK), 0, st, lab, n);
  H_ERR(hipGetLastError());
}

// CC of one graph into lab[n]: the skipped component is the most frequent
// root among `samples` sampled vertices, found on the device (directed
// graphs skip nothing: an out-edge is seen from one end only). `ids` holds
// samples + 1 ints of scratch (the sample, then the chosen root). No host
// read and no allocation: every launch is queued on `st`.
void cc_sampled_launch(const gsw_graph *G, int32_t *lab, int32_t *ids, int samples,
                       hipStream_t st) {
  using namespace gsw;
  if (G->hg.n == 0) return;
  cc_sample_phase(G, lab, st);
  if (G->hg.directed) {
    cc_finish_phase(G, lab, -2, nullptr, nullptr, 1, st);
    return;
  }
  int32_t *skip = ids + samples;
  hipLaunchKernelGGL(k_cc_sample_ids, dim3((samples + BLOCK - 1) / BLOCK), dim3(BLOCK),
                     0, st, ids, samples, G->hg.n);
  hipLaunchKernelGGL(k_cc_mode, dim3(1), dim3(CC_MODE_BLOCK), 0, st,
                     (const int *)lab, (const int32_t *)ids, samples, skip);
  H_ERR(hipGetLastError());
  cc_finish_phase(G, lab, -2, skip, nullptr, 1, st);
}

void cc_sampled_run(gsw_graph *G) {
  cc_buffers(G);
  cc_sampled_launch(G, G->cc_lab, G->cc_ids, CC_SAMPLE, gsw::stream());
}

// Whether no edge of the graph leaves its instance (G->inst_off, count + 1
// offsets): rows are sorted, so one look at each row's first and last
// column decides. Cached per offsets.
bool batch_disjoint(gsw_graph *G, const int64_t *offsets, int64_t count) {
  uint64_t key = 1469598103934665603ull ^ (uint64_t)count;
  for (int64_t i = 0; i <= count; ++i) key = (key ^ (uint64_t)offsets[i]) * 1099511628211ull;
  if (G->disj_known && G->disj_key == key) return G->disj;
  using namespace gsw;
  hipStream_t st = stream();
  set_one_async<int32_t>(G->d_flag, 0);
  hipLaunchKernelGGL(k_cc_cross, dim3(grid_for((G->hg.n + 63) / 64 * 64)), dim3(BLOCK), 0,
                     st, (const int64_t *)G->dg.dg_row_ptr,
                     (const int32_t *)G->dg.dg_adj_list, G->hg.n,
                     (const int64_t *)G->inst_off, count, G->d_flag);
  H_ERR(hipGetLastError());
  H_ERR(hipMemcpyAsync(G->h_flag, G->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  sync();
  G->disj = G->h_flag[0] == 0;
  G->disj_key = key;
  G->disj_known = true;
  return G->disj;
}

// Batched CC over instances no edge leaves, from the BFS levels of every
// instance's first vertex (k_cc_seed_levels): the reached vertices are
// labelled at once and skipped; the others are hooked (two sampled rounds,
// then their remaining edges) and compressed.
void cc_batched_from_levels(gsw_graph *G, int64_t count, const int32_t *lv) {
  using namespace gsw;
  cc_buffers(G);
  hipStream_t st = stream();
  const int64_t n = G->hg.n;
  int32_t *lab = G->cc_lab;
  hipLaunchKernelGGL(k_cc_seed_levels, dim3(grid_for((n + 63) / 64 * 64)), dim3(BLOCK), 0,
                     st, lab, (const int *)lv, (const int64_t *)G->inst_off, count, n);
  for (int r = 0; r < 2; ++r) {
    hipLaunchKernelGGL(k_cc_sample, dim3(grid_for(n)), dim3(BLOCK), 0, st,
                       (const int64_t *)G->dg.dg_row_ptr,
                       (const int32_t *)G->dg.dg_adj_list, n, r, lab, (const int *)lv);
    hipLaunchKernelGGL(k_cc_compress, dim3(grid_for(n)), dim3(BLOCK), 0, st, lab, n);
  }
  hipLaunchKernelGGL(k_cc_rest, dim3(grid_for((n + 63) / 64 * 64)), dim3(BLOCK), 0, st,
                     (const int64_t *)G->dg.dg_row_ptr, (const int32_t *)G->dg.dg_adj_list,
                     n, 2, -1, (const int *)nullptr, (const int64_t *)G->inst_off, count,
                     lab, 1);
  hipLaunchKernelGGL(k_cc_compress, dim3(grid_for(n)), dim3(BLOCK), 0, st, lab, n);
  H_ERR(hipGetLastError());
}

// Batched CC over disjoint instances (offsets on the device): the whole
// kernel sequence is one captured graph, replayed without host decisions.
void cc_batched_captured(gsw_graph *G, int64_t count) {
  cc_buffers(G);
  hipStream_t st = gsw::stream();
  // the sequence's arguments change only with the instance count (a new
  // offsets buffer drops the capture when it is allocated)
  const int64_t key = count;
  if (!G->cc_exec || G->cc_key != key) {
    G->drop_cc_capture();
    H_ERR(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      cc_sample_phase(G, G->cc_lab, st);
      cc_finish_phase(G, G->cc_lab, -1, nullptr, G->inst_off, count, st);
    } catch (...) {
      hipGraph_t dead = nullptr;
      (void)hipStreamEndCapture(st, &dead);
      if (dead) (void)hipGraphDestroy(dead);
      throw;
    }
    H_ERR(hipStreamEndCapture(st, &G->cc_graph));
    H_ERR(hipGraphInstantiate(&G->cc_exec, G->cc_graph, nullptr, nullptr, 0));
    G->cc_key = key;
  }
  H_ERR(hipGraphLaunch(G->cc_exec, st));
}

}  // namespace

extern "C" {

const char *gsw_version(void) { return "gswitch_amd 0.1 (gfx950)"; }

const char *gsw_last_error(void) { return g_last_error.c_str(); }

int32_t gsw_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void gsw_run_opts_init(gsw_run_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->max_levels = 1 << 20;
  o->twc_small = 32;
  o->twc_large = 2048;
  o->profile = 1;
  o->pull_kernel = 2;
}

// ---- host ingest -----------------------------------------------------------
gsw_status gsw_host_graph_from_edges(int64_t n, int64_t ne, const int32_t *src,
                                     const int32_t *dst, const uint32_t *w,
                                     int32_t directed, gsw_host_graph **out) {
  return guarded([&] {
    GSW_REQUIRE(out && (ne == 0 || (src && dst)) && n >= 0 && ne >= 0,
                GSW_E_INVALID, "invalid arguments");
    auto *h = new gsw_host_graph();
    try {
      gsw_host::build_csr(n, ne, src, dst, w, directed != 0, 0, 64, h->g);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

static gsw_status finish_gen(int64_t n, std::vector<int32_t> &s,
                             std::vector<int32_t> &d, uint64_t seed,
                             int32_t weighted, gsw_host_graph **out) {
  auto *h = new gsw_host_graph();
  try {
    const uint64_t wkey =
        weighted ? gsw_host::stream_key(seed, gsw_host::STREAM_WEIGHT) : 0;
    gsw_host::build_csr(n, (int64_t)s.size(), s.data(), d.data(), nullptr,
                        false, wkey, 64, h->g);
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  return GSW_OK;
}

gsw_status gsw_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                        double c, uint64_t seed, int32_t weighted,
                        gsw_host_graph **out) {
  return guarded([&] {
    GSW_REQUIRE(out && scale >= 1 && scale <= 30 && edge_factor >= 1,
                GSW_E_INVALID, "invalid R-MAT parameters");
    GSW_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, GSW_E_INVALID,
                "invalid R-MAT probabilities");
    std::vector<int32_t> s, d;
    gsw_host::gen_rmat(scale, edge_factor, a, b, c, seed, s, d);
    finish_gen((int64_t)1 << scale, s, d, seed, weighted, out);
  });
}

gsw_status gsw_gen_grid2d(int64_t rows, int64_t cols, double p_delete,
                          double long_range_frac, uint64_t seed,
                          int32_t weighted, gsw_host_graph **out) {
  return guarded([&] {
    GSW_REQUIRE(out && rows >= 1 && cols >= 1 && rows * cols <= INT32_MAX,
                GSW_E_INVALID, "invalid grid size");
    std::vector<int32_t> s, d;
    gsw_host::gen_grid2d(rows, cols, p_delete, long_range_frac, seed, s, d);
    finish_gen(rows * cols, s, d, seed, weighted, out);
  });
}

gsw_status gsw_host_graph_save(const gsw_host_graph *hg, const char *path) {
  return guarded([&] {
    GSW_REQUIRE(hg && path, GSW_E_INVALID, "null argument");
    gsw_host::save_csr(hg->g, path);
  });
}

gsw_status gsw_host_graph_load(const char *path, gsw_host_graph **out) {
  return guarded([&] {
    GSW_REQUIRE(out && path, GSW_E_INVALID, "null argument");
    auto *h = new gsw_host_graph();
    try {
      gsw_host::load_csr(path, h->g);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

gsw_status gsw_record_features(const char *path) {
  return guarded([&] {
    record_sink_t &k = record_sink();
    k.path = path ? path : "";
    k.written = 0;
  });
}

gsw_status gsw_read_mtx(const char *path, int32_t directed, int32_t with_weight,
                        int32_t with_header, int32_t rinse,
                        gsw_host_graph **out) {
  return guarded([&] {
    GSW_REQUIRE(out && path, GSW_E_INVALID, "null argument");
    const bool r = rinse < 0 ? (!directed && !with_weight) : rinse != 0;
    auto *h = new gsw_host_graph();
    try {
      gsw_host::read_mtx(path, directed != 0, with_weight != 0, with_header != 0, r,
                         h->g);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

gsw_status gsw_gen_er(int64_t n, double mean_degree, uint64_t seed,
                      int32_t weighted, gsw_host_graph **out) {
  return guarded([&] {
    GSW_REQUIRE(out && n >= 1 && n <= INT32_MAX && mean_degree >= 0,
                GSW_E_INVALID, "invalid ER parameters");
    std::vector<int32_t> s, d;
    gsw_host::gen_er(n, mean_degree, seed, s, d);
    finish_gen(n, s, d, seed, weighted, out);
  });
}

gsw_status gsw_gen_ba(int64_t n, int32_t m, uint64_t seed, int32_t weighted,
                      gsw_host_graph **out) {
  return guarded([&] {
    GSW_REQUIRE(out && n >= 1 && n <= INT32_MAX && m >= 1, GSW_E_INVALID,
                "invalid BA parameters");
    std::vector<int32_t> s, d;
    gsw_host::gen_ba(n, m, seed, s, d);
    finish_gen(n, s, d, seed, weighted, out);
  });
}

gsw_status gsw_gen_er_batch_range(int64_t first, int32_t count, int64_t nmin,
                                  int64_t nmax, double dmin, double dmax,
                                  uint64_t seed, int32_t weighted,
                                  gsw_host_graph **out, int64_t *offsets_out) {
  return guarded([&] {
    GSW_REQUIRE(out && offsets_out && first >= 0 && count >= 1 && nmin >= 1 &&
                    nmax >= nmin && nmax <= INT32_MAX && dmin >= 0 && dmax >= dmin,
                GSW_E_INVALID, "invalid ER batch parameters");
    std::vector<int32_t> s, d;
    std::vector<int64_t> off;
    gsw_host::gen_er_batch(first, count, nmin, nmax, dmin, dmax, seed, s, d, off);
    finish_gen(off.back(), s, d, seed, weighted, out);
    std::memcpy(offsets_out, off.data(), sizeof(int64_t) * off.size());
  });
}

gsw_status gsw_gen_er_batch(int32_t count, int64_t nmin, int64_t nmax,
                            double dmin, double dmax, uint64_t seed,
                            int32_t weighted, gsw_host_graph **out,
                            int64_t *offsets_out) {
  return gsw_gen_er_batch_range(0, count, nmin, nmax, dmin, dmax, seed,
                                weighted, out, offsets_out);
}

gsw_status gsw_host_graph_info(const gsw_host_graph *g, int64_t *n, int64_t *m,
                               int32_t *directed, int32_t *weighted) {
  return guarded([&] {
    GSW_REQUIRE(g, GSW_E_INVALID, "null graph");
    if (n) *n = g->g.n;
    if (m) *m = g->g.m;
    if (directed) *directed = g->g.directed;
    if (weighted) *weighted = g->g.weighted;
  });
}

gsw_status gsw_host_graph_export(const gsw_host_graph *g, int64_t *row_ptr,
                                 int32_t *col, uint32_t *w) {
  return guarded([&] {
    GSW_REQUIRE(g && row_ptr && (col || g->g.m == 0), GSW_E_INVALID,
                "invalid arguments");
    std::memcpy(row_ptr, g->g.row_ptr.data(), sizeof(int64_t) * (g->g.n + 1));
    if (g->g.m) std::memcpy(col, g->g.col.data(), sizeof(int32_t) * g->g.m);
    if (w && g->g.weighted && g->g.m)
      std::memcpy(w, g->g.w.data(), sizeof(uint32_t) * g->g.m);
  });
}

void gsw_host_graph_destroy(gsw_host_graph *g) { delete g; }

gsw_status gsw_degree_stats(const int64_t *row_ptr, int64_t n, double *out6) {
  return guarded([&] {
    GSW_REQUIRE(row_ptr && out6 && n > 0, GSW_E_INVALID, "invalid arguments");
    graph_attr_t a;
    a.fill(row_ptr, n);
    out6[0] = a.avg_deg;
    out6[1] = a.std_deg;
    out6[2] = a.max_deg;
    out6[3] = a.range;
    out6[4] = a.GI;
    out6[5] = a.Her;
  });
}

// ---- device graph ----------------------------------------------------------
gsw_status gsw_graph_create_csr(const int64_t *row_ptr, const int32_t *col,
                                const uint32_t *w, int64_t n, int64_t m,
                                int32_t directed, gsw_graph **out) {
  return guarded([&] {
    GSW_REQUIRE(out && row_ptr && (col || m == 0) && n >= 1 && m >= 0 &&
                    n <= INT32_MAX,
                GSW_E_INVALID, "invalid arguments");
    GSW_REQUIRE(!directed, GSW_E_UNSUPPORTED,
                "gsw_graph_create_csr takes symmetric graphs; use "
                "gsw_host_graph_from_edges for directed input");
    GSW_REQUIRE(row_ptr[0] == 0 && row_ptr[n] == m, GSW_E_INVALID,
                "row_ptr must start at 0 and end at m");
    for (int64_t v = 0; v < n; ++v)
      GSW_REQUIRE(row_ptr[v] <= row_ptr[v + 1], GSW_E_INVALID,
                  "row_ptr must be non-decreasing");
    for (int64_t j = 0; j < m; ++j)
      GSW_REQUIRE(col[j] >= 0 && col[j] < n, GSW_E_INVALID,
                  "column index out of range");
    require_gpu();
    auto G = std::make_unique<gsw_graph>();
    G->hg.n = n;
    G->hg.m = m;
    G->hg.directed = false;
    G->hg.weighted = (w != nullptr);
    G->hg.row_ptr.assign(row_ptr, row_ptr + n + 1);
    G->hg.col.assign(col, col + m);
    if (w) G->hg.w.assign(w, w + m);
    sort_rows(G->hg);
    try {
      upload_graph(G.get());
    } catch (...) {
      G->release();
      throw;
    }
    *out = G.release();
  });
}

gsw_status gsw_graph_from_host(const gsw_host_graph *hg, gsw_graph **out) {
  return guarded([&] {
    GSW_REQUIRE(out && hg && hg->g.n >= 1, GSW_E_INVALID, "invalid arguments");
    require_gpu();
    auto G = std::make_unique<gsw_graph>();
    G->hg = hg->g;
    try {
      upload_graph(G.get());
    } catch (...) {
      G->release();
      throw;
    }
    *out = G.release();
  });
}

gsw_status gsw_graph_destroy(gsw_graph *g) {
  return guarded([&] {
    if (!g) return;
    gsw::sync();
    g->release();
    delete g;
  });
}

gsw_status gsw_last_records(const gsw_graph *g, gsw_level_record *out,
                            int32_t cap, int32_t *n) {
  return guarded([&] {
    GSW_REQUIRE(g && n && (out || cap == 0), GSW_E_INVALID, "invalid arguments");
    const int32_t k = (int32_t)MIN((int64_t)cap, (int64_t)g->records.size());
    if (k) std::memcpy(out, g->records.data(), sizeof(gsw_level_record) * k);
    *n = k;
  });
}

// ---- BFS -------------------------------------------------------------------
gsw_status gsw_bfs(gsw_graph *G, int32_t root, const gsw_run_opts *opts,
                   int32_t *levels_out, gsw_run_stats *stats) {
  return guarded([&] {
    GSW_REQUIRE(G, GSW_E_INVALID, "null graph");
    GSW_REQUIRE(root >= 0 && root < G->hg.n, GSW_E_INVALID, "root out of range");
    int levels = 0;
    if (opts && opts->capture > 0) {
      GSW_REQUIRE(!opts->configs && opts->selector == 0, GSW_E_UNSUPPORTED,
                  "the device-resident BFS picks its own directions "
                  "(no configs or selector)");
      const run_opts_t o = to_opts(opts);
      const int grid = opts->grid > 0 ? opts->grid : default_do_grid(G);
      levels = bfs_device_run(G, root, opts->capture, grid, o.max_levels);
    } else {
      levels = bfs_engine_run(G, opts, root, nullptr, 1, 0.0);
      collect_records(G, G->bfs->en.executor, G->bfs->as);
    }
    const double ms = elapsed(G);
    auto &c = *G->bfs;
    if (levels_out)
      H_ERR(hipMemcpy(levels_out, c.f.data.dg_wa, sizeof(int32_t) * G->hg.n,
                      hipMemcpyDeviceToHost));
    if (stats) {
      stats->device_ms = ms;
      stats->levels = levels;
      stats->n_records = (int32_t)G->records.size();
      stats->reached = 0;
    }
  });
}

// ---- batched / multi-root BFS ---------------------------------------------
gsw_status gsw_bfs_multi(gsw_graph *G, const int32_t *roots, int32_t nroots,
                         const gsw_run_opts *opts, int32_t *levels_out,
                         gsw_run_stats *stats) {
  return guarded([&] {
    GSW_REQUIRE(G, GSW_E_INVALID, "null graph");
    double workload = 0;
    const std::vector<int32_t> r = upload_roots(G, roots, nroots, workload);
    const int64_t k = (int64_t)r.size();
    int levels = 0;
    if (opts && opts->capture > 0) {
      const run_opts_t o = to_opts(opts);
      LB lb = WM;
      if (o.force.has_lb) {
        GSW_REQUIRE(o.force.lb == WM || o.force.lb == TWC, GSW_E_UNSUPPORTED,
                    "captured BFS supports the WM and TWC expands");
        lb = o.force.lb;
      }
      const int grid = opts->grid > 0 ? opts->grid : default_capture_grid(G);
      levels = bfs_captured_run(G, k, lb, o.twc_small, o.twc_large, grid,
                                2 * opts->capture, o.max_levels);
      G->records.clear();
    } else {
      levels = bfs_engine_run(G, opts, r[0], G->d_roots, k, workload);
      collect_records(G, G->bfs->en.executor, G->bfs->as);
    }
    const double ms = elapsed(G);
    if (levels_out)
      H_ERR(hipMemcpy(levels_out, G->bfs->f.data.dg_wa,
                      sizeof(int32_t) * G->hg.n, hipMemcpyDeviceToHost));
    if (stats) {
      stats->device_ms = ms;
      stats->levels = levels;
      stats->n_records = (int32_t)G->records.size();
      stats->reached = 0;
    }
  });
}

gsw_status gsw_batch_bfs_cc(gsw_graph *G, const int64_t *offsets, int32_t count,
                            const gsw_run_opts *opts, int32_t *levels_out,
                            int32_t *labels_out, gsw_run_stats *stats) {
  return guarded([&] {
    GSW_REQUIRE(G && offsets && count >= 1, GSW_E_INVALID, "invalid arguments");
    GSW_REQUIRE(!G->hg.directed, GSW_E_UNSUPPORTED,
                "batched BFS + CC needs an undirected graph");
    GSW_REQUIRE(offsets[0] == 0 && offsets[count] == G->hg.n, GSW_E_INVALID,
                "instance offsets must run from 0 to n");
    std::vector<int32_t> roots;
    for (int32_t i = 0; i < count; ++i) {
      GSW_REQUIRE(offsets[i] <= offsets[i + 1], GSW_E_INVALID,
                  "instance offsets must not decrease");
      if (offsets[i] < offsets[i + 1]) roots.push_back((int32_t)offsets[i]);
    }
    const run_opts_t o = to_opts(opts);
    LB lb = WM;
    if (o.force.has_lb) {
      GSW_REQUIRE(o.force.lb == WM || o.force.lb == TWC, GSW_E_UNSUPPORTED,
                  "captured BFS supports the WM and TWC expands");
      lb = o.force.lb;
    }
    const int capture = (opts && opts->capture > 0) ? opts->capture : 8;
    const int grid = (opts && opts->grid > 0) ? opts->grid : default_capture_grid(G);
    double workload = 0;
    upload_roots(G, roots.data(), (int32_t)roots.size(), workload);
    if (count + 1 > G->inst_cap) {
      if (G->inst_off) H_ERR(hipFree(G->inst_off));
      G->inst_off = nullptr;
      H_ERR(hipMalloc((void **)&G->inst_off, sizeof(int64_t) * (count + 1)));
      G->inst_cap = count + 1;
      G->drop_cc_capture();
    }
    H_ERR(hipMemcpy(G->inst_off, offsets, sizeof(int64_t) * (count + 1),
                    hipMemcpyHostToDevice));
    // BFS from every instance's first vertex: a forced push runs the
    // captured fused level loop; otherwise the device-resident
    // direction-optimising loop from all the roots at once
    int levels;
    if (o.force.has_direction) {
      GSW_REQUIRE(o.force.d == Push, GSW_E_UNSUPPORTED,
                  "the captured batch loop pushes; leave the direction free for the "
                  "direction-optimising loop");
      levels = bfs_captured_run(G, (int64_t)roots.size(), lb, o.twc_small, o.twc_large,
                                grid, 2 * capture, o.max_levels);
    } else {
      const int dgrid = (opts && opts->grid > 0) ? opts->grid : default_do_grid(G);
      levels = bfs_device_run(G, -1, capture, dgrid, o.max_levels,
                              (int64_t)roots.size());
    }
    // CC: from the BFS levels when no edge leaves an instance, otherwise one
    // captured launch sequence over the same graph
    if (batch_disjoint(G, offsets, count))
      cc_batched_from_levels(G, count, G->bfs->f.data.dg_wa);
    else
      cc_batched_captured(G, count);
    H_ERR(hipEventRecord(G->ev1, gsw::stream()));
    const double ms = elapsed(G);
    G->records.clear();
    if (levels_out)
      H_ERR(hipMemcpy(levels_out, G->bfs->f.data.dg_wa, sizeof(int32_t) * G->hg.n,
                      hipMemcpyDeviceToHost));
    if (labels_out)
      H_ERR(hipMemcpy(labels_out, G->cc_lab, sizeof(int32_t) * G->hg.n,
                      hipMemcpyDeviceToHost));
    if (stats) {
      stats->device_ms = ms;
      stats->levels = levels;
      stats->n_records = 0;
      stats->reached = 0;
    }
  });
}

gsw_status gsw_autotune_bfs(gsw_graph *G, const int32_t *roots, int32_t nroots,
                            int32_t capture, int32_t *best_lb,
                            int32_t *best_small, int32_t *best_large,
                            int32_t *best_grid, double *best_ms) {
  return guarded([&] {
    GSW_REQUIRE(G && capture > 0 && best_lb && best_small && best_large &&
                    best_grid && best_ms,
                GSW_E_INVALID, "invalid arguments");
    double workload = 0;
    const std::vector<int32_t> r = upload_roots(G, roots, nroots, workload);
    const int64_t k = (int64_t)r.size();
    struct cand_t { LB lb; int small, large; };
    const cand_t cands[] = {{WM, 32, 2048},   {TWC, 16, 1024}, {TWC, 32, 2048},
                            {TWC, 64, 4096},  {TWC, 128, 8192}};
    const int gmax = default_capture_grid(G);
    int grids[3] = {MAX(1, gmax / 8), MAX(1, gmax / 2), gmax};
    double best = 1e300;
    for (const cand_t &cd : cands)
      for (int gi = 0; gi < 3; ++gi) {
        if (gi > 0 && grids[gi] == grids[gi - 1]) continue;
        double t = 1e300;
        for (int rep = 0; rep < 3; ++rep) {
          bfs_captured_run(G, k, cd.lb, cd.small, cd.large, grids[gi],
                           2 * capture, 1 << 20);
          t = std::min(t, elapsed(G));
        }
        if (t < best) {
          best = t;
          *best_lb = (int32_t)cd.lb;
          *best_small = cd.small;
          *best_large = cd.large;
          *best_grid = grids[gi];
        }
      }
    *best_ms = best;
  });
}

// ---- SSSP stepping window per graph --------------------------------------
gsw_status gsw_sssp_set_window(gsw_graph *G, double scale) {
  return guarded([&] {
    GSW_REQUIRE(G && scale > 0 && scale <= 1024, GSW_E_INVALID, "invalid window scale");
    G->sssp_window_scale = scale;
  });
}

gsw_status gsw_sssp_get_window(const gsw_graph *G, double *scale) {
  return guarded([&] {
    GSW_REQUIRE(G && scale, GSW_E_INVALID, "null argument");
    *scale = G->sssp_window_scale;
  });
}

gsw_status gsw_autotune_sssp(gsw_graph *G, const int32_t *roots, int32_t nroots,
                             const gsw_run_opts *opts, double *best_scale,
                             double *best_ms) {
  return guarded([&] {
    GSW_REQUIRE(G && roots && nroots > 0 && best_scale && best_ms, GSW_E_INVALID,
                "invalid arguments");
    GSW_REQUIRE(G->has_w, GSW_E_INVALID, "SSSP needs a weighted graph");
    for (int32_t i = 0; i < nroots; ++i)
      GSW_REQUIRE(roots[i] >= 0 && roots[i] < G->hg.n, GSW_E_INVALID,
                  "root out of range");
    gsw_run_opts o;
    gsw_run_opts_init(&o);
    if (opts) o = *opts;
    o.profile = 0;
    const double cands[] = {1.0, 0.5, 0.25, 0.125, 0.0625};
    double best = 1e300, pick = 1.0;
    // a failed run leaves the graph's scale as it was
    struct keep_t {
      gsw_graph *g;
      double k;
      bool done = false;
      ~keep_t() {
        if (!done) g->sssp_window_scale = k;
      }
    } keep{G, G->sssp_window_scale};
    for (double k : cands) {
      G->sssp_window_scale = k;
      double total = 0;
      for (int32_t i = 0; i < nroots; ++i) {
        double t = 1e300;
        for (int rep = 0; rep < 2; ++rep) {
          gsw_run_stats st{};
          const gsw_status rc = gsw_sssp(G, roots[i], &o, nullptr, &st);
          GSW_REQUIRE(rc == GSW_OK, rc, std::string("SSSP failed: ") + gsw_last_error());
          t = std::min(t, st.device_ms);
        }
        total += t;
      }
      if (total < best) {
        best = total;
        pick = k;
      }
    }
    G->sssp_window_scale = pick;
    keep.done = true;
    *best_scale = pick;
    *best_ms = best;
  });
}

// ---- BC --------------------------------------------------------------------
// run_bc (application/bc.cu:97-160): forward level loop with BCF, then the
// backward loop with BCB from the deepest level up to the root. The backward
// pass always runs standalone levels (the inspector applies the dependency
// finalisation to every vertex of the level, including those no child
// reached), and each level's frontier comes from the inspector's bitmaps.
gsw_status gsw_bc(gsw_graph *G, int32_t root, const gsw_run_opts *opts,
                  int32_t *levels_out, double *nsp_out, double *dep_out,
                  gsw_run_stats *stats) {
  return guarded([&] {
    GSW_REQUIRE(G, GSW_E_INVALID, "null graph");
    GSW_REQUIRE(root >= 0 && root < G->hg.n, GSW_E_INVALID, "root out of range");
    GSW_REQUIRE(!G->hg.directed, GSW_E_UNSUPPORTED,
                "BC needs an undirected graph");
    const run_opts_t o = to_opts_rule(opts);
    const int64_t n = G->hg.n;
    if (!G->bc) {
      G->bc = std::make_unique<app_ctx_t<BCF>>();
      G->bc->as.build(n, Normal);
      G->bc->f.data.build(n);
    }
    if (!G->bcb_built) {
      G->bcb.data.build(n);
      G->bcb_built = true;
    }
    if (!G->d_bc_out)
      H_ERR(hipMalloc((void **)&G->d_bc_out, 20 * (size_t)(n + 1)));
    auto &c = *G->bc;
    BCB &h = G->bcb;
    arch_once(c);
    hipStream_t st = gsw::stream();

    // forward
    prepare_engine(c, o, !opts || opts->profile);
    feature_t &fets = c.en.fets;
    fets.centric = VC;
    fets.pattern = ASSO;
    fets.fromall = false;
    fets.toall = false;
    fets.use_root = root;
    fets.set_graph((double)n, (double)(G->hg.m / 2), G->attr);
    view_t<GU> gv{&G->hg, G->attr, 0.0};
    init_conf(c.en.stats, fets, c.en.conf, gv, c.f);
    H_ERR(hipEventRecord(G->ev0, st));
    c.as.reset();
    c.as.init(root, true);
    hipLaunchKernelGGL(k_init_bc, dim3(gsw::grid_for(n)), dim3(gsw::BLOCK), 0, st,
                       c.f.data.dg_wa, n, root);
    GU g = G->dg;
    int levels = run_vc(c.en, c.as, g, c.f, o, [](int) { return false; });
    collect_records(G, c.en.executor, c.as);
    std::vector<gsw_level_record> fwd_records = G->records;

    // backward
    gsw::set_one_async<int32_t>(G->d_flag, -1);
    hipLaunchKernelGGL(k_bc_backward_init, dim3(gsw::grid_for(n)),
                       dim3(gsw::BLOCK), 0, st, c.f.data.dg_wa, h.data.dg_wa,
                       h.data.dg_ra, n, G->d_flag);
    H_ERR(hipGetLastError());
    H_ERR(hipMemcpyAsync(G->h_flag, G->d_flag, sizeof(int32_t),
                         hipMemcpyDeviceToHost, st));
    gsw::sync();
    const int deepest = G->h_flag[0];
    prepare_engine(c, o, !opts || opts->profile);
    fets.centric = VC;
    fets.pattern = ASSO;
    fets.fromall = false;
    fets.toall = false;
    fets.use_root = -1;
    fets.set_graph((double)n, (double)(G->hg.m / 2), G->attr);
    c.en.conf.conf_fuse_inspect = false;
    c.en.conf.force.has_fusion = true;
    c.en.conf.force.fusion = Standalone;
    c.en.conf.conf_first_round = false;
    c.as.reset();
    c.as.init(ALL_INACTIVE, true);
    g.level = deepest;
    for (int l = 0; l < o.max_levels && g.level >= 0; ++l) {
      c.en.inspector.inspect(c.as, g, h, c.en.stats, fets, c.en.conf);
      if (g.level == 0) break;  // the root is final; it has no parents
      if (c.as.finish(g, h, c.en.conf)) break;
      c.en.selector.select(c.en.stats, fets, c.en.conf);
      c.en.executor.filter(c.as, g, h, c.en.stats, fets, c.en.conf);
      g.update_level(-1);
      c.en.executor.expand(c.as, g, h, c.en.stats, fets, c.en.conf);
      if (c.as.finish(g, h, c.en.conf)) break;
      ++levels;
    }
    H_ERR(hipEventRecord(G->ev1, st));
    const double ms = elapsed(G);
    collect_records(G, c.en.executor, c.as);
    G->records.insert(G->records.begin(), fwd_records.begin(), fwd_records.end());
    if (levels_out || nsp_out || dep_out) {
      double *d_nsp = (double *)G->d_bc_out;
      double *d_dep = d_nsp + n;
      int32_t *d_lvl = (int32_t *)(d_dep + n);
      hipLaunchKernelGGL(k_bc_output, dim3(gsw::grid_for(n)), dim3(gsw::BLOCK),
                         0, st, c.f.data.dg_wa, h.data.dg_wa, n, d_lvl, d_nsp,
                         d_dep);
      H_ERR(hipGetLastError());
      if (levels_out)
        H_ERR(hipMemcpy(levels_out, d_lvl, sizeof(int32_t) * n,
                        hipMemcpyDeviceToHost));
      if (nsp_out)
        H_ERR(hipMemcpy(nsp_out, d_nsp, sizeof(double) * n, hipMemcpyDeviceToHost));
      if (dep_out)
        H_ERR(hipMemcpy(dep_out, d_dep, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    if (stats) {
      stats->device_ms = ms;
      stats->levels = levels;
      stats->n_records = (int32_t)G->records.size();
      stats->reached = 0;
    }
  });
}

// ---- SSSP ------------------------------------------------------------------
gsw_status gsw_sssp(gsw_graph *G, int32_t root, const gsw_run_opts *opts,
                    uint32_t *dist_out, gsw_run_stats *stats) {
  return guarded([&] {
    GSW_REQUIRE(G, GSW_E_INVALID, "null graph");
    GSW_REQUIRE(G->has_w, GSW_E_INVALID, "SSSP needs a weighted graph");
    GSW_REQUIRE(root >= 0 && root < G->hg.n, GSW_E_INVALID, "root out of range");
    const run_opts_t o = to_opts(opts);
    auto &c = sssp_ctx(G);
    arch_once(c);
    sssp_prepare(G, c, o, root, !opts || opts->profile);
    maybe_learned(c, o, G, false, true, GSW_APP_SSSP);

    hipStream_t st = gsw::stream();
    H_ERR(hipEventRecord(G->ev0, st));
    c.as.reset();
    c.as.init(root, true);
    hipLaunchKernelGGL(k_init_sssp, dim3(gsw::grid_for(G->hg.n)),
                       dim3(gsw::BLOCK), 0, st, c.f.data.dg_wa, c.f.data.dg_ra,
                       G->hg.n, root);
    GW g = G->dgw;
    int levels = run_vc(c.en, c.as, g, c.f, o, [](int) { return false; });
    H_ERR(hipEventRecord(G->ev1, st));
    const double ms = elapsed(G);
    collect_records(G, c.en.executor, c.as);
    if (dist_out)
      H_ERR(hipMemcpy(dist_out, c.f.data.dg_wa, sizeof(uint32_t) * G->hg.n,
                      hipMemcpyDeviceToHost));
    if (stats) {
      stats->device_ms = ms;
      stats->levels = levels;
      stats->n_records = (int32_t)G->records.size();
      stats->reached = 0;
    }
  });
}

// ---- CC --------------------------------------------------------------------
gsw_status gsw_cc(gsw_graph *G, const gsw_run_opts *opts, int32_t *labels_out,
                  gsw_run_stats *stats) {
  return guarded([&] {
    GSW_REQUIRE(G, GSW_E_INVALID, "null graph");
    run_opts_t o = to_opts(opts);
    const bool forced = opts && opts->configs;
    int cls = -1;
    if (o.selector != 0 && !forced) cls = choose_once(G, GSW_APP_CC, o.selector,
                                                      gsw_mlp::CC_CLASSES);
    G->last_class = cls;
    if (cls >= 2) {
      o.force.has_asfmt = true;
      o.force.asf = cls == 2 ? Bitmap : Queue;
    }
    if (cls == 0 || (cls < 0 && !forced)) {
      // sampled hooking over the CSR
      hipStream_t st = gsw::stream();
      H_ERR(hipEventRecord(G->ev0, st));
      cc_sampled_run(G);
      H_ERR(hipEventRecord(G->ev1, st));
      const double ms = elapsed(G);
      G->records.clear();
      if (labels_out)
        H_ERR(hipMemcpy(labels_out, G->cc_lab, sizeof(int32_t) * G->hg.n,
                        hipMemcpyDeviceToHost));
      if (stats) {
        stats->device_ms = ms;
        stats->levels = 1;
        stats->n_records = 0;
        stats->reached = 0;
      }
      return;
    }
    if (!G->has_coo) {
      G->coo.build(G->hg);
      G->has_coo = true;
    }
    GSW_REQUIRE(G->coo.nedges <= INT32_MAX, GSW_E_UNSUPPORTED,
                "the edge-centric frontier holds int32 edge ids: CC takes at "
                "most 2^31 - 1 undirected edges");
    if (!G->cc) {
      G->cc = std::make_unique<app_ctx_t<CC>>();
      G->cc->as.build(MAX((int64_t)1, G->coo.nedges), Normal);
      G->cc->f.data.build(G->hg.n);
    }
    auto &c = *G->cc;
    arch_once(c);
    prepare_engine(c, o, !opts || opts->profile);
    feature_t &fets = c.en.fets;
    fets.centric = EC;
    fets.pattern = ASSO;
    fets.set_graph((double)G->hg.n, (double)G->coo.nedges, G->attr);
    view_t<GC> gv{&G->hg, G->attr, 0.0};
    init_conf(c.en.stats, fets, c.en.conf, gv, c.f);

    hipStream_t st = gsw::stream();
    H_ERR(hipEventRecord(G->ev0, st));
    c.as.reset();
    c.as.init(ALL_ACTIVE);
    hipLaunchKernelGGL(k_init_iota, dim3(gsw::grid_for(G->hg.n)),
                       dim3(gsw::BLOCK), 0, st, c.f.data.dg_wa, G->hg.n);
    GC g = G->coo;
    int level = 0;
    if (g.nedges > 0) {
      for (level = 0; level < o.max_levels; ++level) {
        c.en.inspector.inspect(c.as, g, c.f, c.en.stats, fets, c.en.conf);
        if (c.as.finish(g, c.f, c.en.conf)) break;
        c.en.selector.select(c.en.stats, fets, c.en.conf);
        c.en.executor.filter(c.as, g, c.f, c.en.stats, fets, c.en.conf);
        g.update_level();
        c.en.executor.expand(c.as, g, c.f, c.en.stats, fets, c.en.conf);
        p_jump(c.f, G->hg.n, G->d_flag, G->h_flag);
        if (c.as.finish(g, c.f, c.en.conf)) break;
      }
    }
    H_ERR(hipEventRecord(G->ev1, st));
    const double ms = elapsed(G);
    collect_records(G, c.en.executor, c.as);
    if (labels_out)
      H_ERR(hipMemcpy(labels_out, c.f.data.dg_wa, sizeof(int32_t) * G->hg.n,
                      hipMemcpyDeviceToHost));
    if (stats) {
      stats->device_ms = ms;
      stats->levels = level;
      stats->n_records = (int32_t)G->records.size();
      stats->reached = 0;
    }
  });
}

// ---- PageRank --------------------------------------------------------------
gsw_status gsw_pagerank(gsw_graph *G, float damping, int32_t iters, float eps,
                        const gsw_run_opts *opts, float *pr_out,
                        gsw_run_stats *stats) {
  return guarded([&] {
    GSW_REQUIRE(G, GSW_E_INVALID, "null graph");
    GSW_REQUIRE(iters >= 0 && damping >= 0 && damping <= 1 && eps >= 0,
                GSW_E_INVALID, "invalid PageRank parameters");
    const run_opts_t o = to_opts(opts);
    if (!G->pr) {
      G->pr = std::make_unique<app_ctx_t<PR>>();
      G->pr->as.build(G->hg.n, Normal);
      G->pr->f.data.build(G->hg.n);
    }
    auto &c = *G->pr;
    arch_once(c);
    prepare_engine(c, o, !opts || opts->profile);
    c.f.damp = damping;
    c.f.eps = eps;
    feature_t &fets = c.en.fets;
    fets.centric = VC;
    fets.pattern = ASSO;
    fets.fromall = true;
    fets.toall = true;
    fets.set_graph((double)G->hg.n, (double)(G->hg.m / 2), G->attr);
    view_t<GU> gv{&G->hg, G->attr, 0.0};
    init_conf(c.en.stats, fets, c.en.conf, gv, c.f);
    maybe_learned(c, o, G, true, true, GSW_APP_PR);
    // with eps = 0 the filter keeps every vertex active: no per-level reads
    c.en.conf.counts_known = eps == 0.f;
    // PR's filter zeroes the accumulator it consumes (GSW_PB_WA_ZERO=0 reads
    // it back anyway)
    {
      const char *e = std::getenv("GSW_PB_WA_ZERO");
      c.en.conf.wa_zero = !(e && std::atoi(e) == 0);
    }

    hipStream_t st = gsw::stream();
    H_ERR(hipEventRecord(G->ev0, st));
    c.as.reset();
    c.as.init(ALL_ACTIVE);
    hipLaunchKernelGGL(k_init_pr, dim3(gsw::grid_for(G->hg.n)), dim3(gsw::BLOCK),
                       0, st, c.f.data.dg_wa, c.f.data.dg_ra, G->dg.dg_row_ptr,
                       G->hg.n, damping);
    GU g = G->dg;
    const int cap = iters;
    int levels = run_vc(c.en, c.as, g, c.f, o,
                        [cap](int level) { return level >= cap; });
    hipLaunchKernelGGL(k_pr_output, dim3(gsw::grid_for(G->hg.n)),
                       dim3(gsw::BLOCK), 0, st, c.f.data.dg_ra,
                       G->dg.dg_row_ptr, G->hg.n, G->d_out_f);
    H_ERR(hipEventRecord(G->ev1, st));
    const double ms = elapsed(G);
    collect_records(G, c.en.executor, c.as);
    G->last_class = o.selector != 0 ? c.en.conf.mlp_applied : -1;
    if (pr_out)
      H_ERR(hipMemcpy(pr_out, G->d_out_f, sizeof(float) * G->hg.n,
                      hipMemcpyDeviceToHost));
    if (stats) {
      stats->device_ms = ms;
      stats->levels = levels;
      stats->n_records = (int32_t)G->records.size();
      stats->reached = 0;
    }
  });
}

// Delta-PageRank: every vertex starts active with value 1 - d and sends it;
// the run ends when no accumulation exceeds eps (or after max_iters rounds).
gsw_status gsw_delta_pagerank(gsw_graph *G, float damping, double eps,
                              int32_t max_iters, const gsw_run_opts *opts,
                              float *pr_out, gsw_run_stats *stats) {
  return guarded([&] {
    GSW_REQUIRE(G, GSW_E_INVALID, "null graph");
    GSW_REQUIRE(damping > 0 && damping < 1 && eps > 0 && max_iters >= 1, GSW_E_INVALID,
                "Delta-PageRank needs 0 < damping < 1, eps > 0, max_iters >= 1");
    const run_opts_t o = to_opts_rule(opts);
    if (!G->dpr) {
      G->dpr = std::make_unique<app_ctx_t<DeltaPR>>();
      G->dpr->as.build(G->hg.n, Normal);
      G->dpr->f.data.build(G->hg.n);
    }
    auto &c = *G->dpr;
    arch_once(c);
    prepare_engine(c, o, !opts || opts->profile);
    c.f.damp = damping;
    c.f.eps = eps;
    feature_t &fets = c.en.fets;
    fets.centric = VC;
    fets.pattern = ASSO;
    fets.fromall = false;
    fets.toall = true;
    fets.set_graph((double)G->hg.n, (double)(G->hg.m / 2), G->attr);
    view_t<GU> gv{&G->hg, G->attr, 0.0};
    init_conf(c.en.stats, fets, c.en.conf, gv, c.f);
    // the filter applies the accumulated changes, so every level must be
    // inspected: no fused levels (a fused level would skip the filter)
    GSW_REQUIRE(!(c.en.conf.force.has_fusion && c.en.conf.force.fusion == Fused),
                GSW_E_INVALID, "Delta-PageRank runs standalone levels only");
    c.en.conf.conf_fuse_inspect = false;
    c.en.conf.conf_fusion = false;

    hipStream_t st = gsw::stream();
    H_ERR(hipEventRecord(G->ev0, st));
    c.as.reset();
    c.as.init(ALL_ACTIVE);
    hipLaunchKernelGGL(k_init_dpr, dim3(gsw::grid_for(G->hg.n)), dim3(gsw::BLOCK),
                       0, st, c.f.data.dg_wa, c.f.data.dg_ra, G->dg.dg_row_ptr,
                       G->hg.n, damping);
    GU g = G->dg;
    const int cap = max_iters;
    const int levels = run_vc(c.en, c.as, g, c.f, o,
                              [cap](int level) { return level >= cap; });
    hipLaunchKernelGGL(k_dpr_output, dim3(gsw::grid_for(G->hg.n)), dim3(gsw::BLOCK),
                       0, st, c.f.data.dg_ra, G->hg.n, G->d_out_f);
    H_ERR(hipEventRecord(G->ev1, st));
    const double ms = elapsed(G);
    collect_records(G, c.en.executor, c.as);
    if (pr_out)
      H_ERR(hipMemcpy(pr_out, G->d_out_f, sizeof(float) * G->hg.n,
                      hipMemcpyDeviceToHost));
    if (stats) {
      stats->device_ms = ms;
      stats->levels = levels;
      stats->n_records = (int32_t)G->records.size();
      stats->reached = 0;
    }
  });
}

// run_gc (application/gc.cu:76-98, main :100-127): from-all, to-all,
// associative, Fixed vertices leave the frontier for good.
gsw_status gsw_gc(gsw_graph *G, const gsw_run_opts *opts, int32_t *colors_out,
                  gsw_run_stats *stats) {
  return guarded([&] {
    GSW_REQUIRE(G, GSW_E_INVALID, "null graph");
    // Jones-Plassmann over one edge direction would let both ends of a
    // one-way edge take the same colour
    GSW_REQUIRE(!G->hg.directed, GSW_E_UNSUPPORTED,
                "graph colouring needs an undirected graph");
    const run_opts_t o = to_opts_rule(opts);
    if (!G->gc) {
      G->gc = std::make_unique<app_ctx_t<Color>>();
      G->gc->as.build(G->hg.n, Normal);
      G->gc->f.data.build(G->hg.n);
    }
    auto &c = *G->gc;
    arch_once(c);
    prepare_engine(c, o, !opts || opts->profile);
    feature_t &fets = c.en.fets;
    fets.centric = VC;
    fets.pattern = ASSO;
    fets.fromall = true;
    fets.toall = true;
    c.en.conf.conf_fixed = true;
    fets.set_graph((double)G->hg.n, (double)(G->hg.m / 2), G->attr);
    view_t<GU> gv{&G->hg, G->attr, 0.0};
    init_conf(c.en.stats, fets, c.en.conf, gv, c.f);

    hipStream_t st = gsw::stream();
    H_ERR(hipEventRecord(G->ev0, st));
    c.as.reset();
    c.as.init(ALL_ACTIVE);
    gsw::fill_async<int32_t>(c.f.data.dg_wa, -1, G->hg.n);
    gsw::fill_async<int32_t>(c.f.data.dg_ra, -1, G->hg.n);
    GU g = G->dg;
    const int levels = run_vc(c.en, c.as, g, c.f, o, [](int) { return false; });
    H_ERR(hipEventRecord(G->ev1, st));
    const double ms = elapsed(G);
    collect_records(G, c.en.executor, c.as);
    if (colors_out)
      H_ERR(hipMemcpy(colors_out, c.f.data.dg_ra, sizeof(int32_t) * G->hg.n,
                      hipMemcpyDeviceToHost));
    if (stats) {
      stats->device_ms = ms;
      stats->levels = levels;
      stats->n_records = (int32_t)G->records.size();
      stats->reached = 0;
    }
  });
}

// ---- selector trees ----------------------------------------------------------
// ---- 64-source bitmask BFS ------------------------------------------------
gsw_status gsw_msbfs64(gsw_graph *G, const int32_t *sources, int32_t nsrc,
                       const gsw_run_opts *opts, uint64_t *mask_out,
                       uint64_t *level_sums_out, gsw_run_stats *stats) {
  return guarded([&] {
    using namespace gsw;
    GSW_REQUIRE(G && sources && nsrc >= 1 && nsrc <= 64, GSW_E_INVALID,
                "need 1..64 sources");
    const int64_t n = G->hg.n, m = G->hg.m;
    for (int s = 0; s < nsrc; ++s)
      GSW_REQUIRE(sources[s] >= 0 && sources[s] < n, GSW_E_INVALID,
                  "source out of range");
    const run_opts_t o = to_opts(opts);
    if (!G->ms_buf) {
      H_ERR(hipMalloc((void **)&G->ms_buf, sizeof(uint64_t) * 3 * (n + 1)));
      H_ERR(hipMalloc((void **)&G->ms_cnt, sizeof(ms_counters_t) * MS_SLOTS));
      H_ERR(hipHostMalloc((void **)&G->ms_hcnt, sizeof(ms_counters_t) * MS_SLOTS));
      H_ERR(hipMalloc((void **)&G->ms_src, sizeof(int32_t) * 64));
    }
    uint64_t *visit = G->ms_buf, *front = G->ms_buf + (n + 1),
             *next = G->ms_buf + 2 * (n + 1);
    const uint64_t all = (nsrc == 64) ? ~0ull : ((1ull << nsrc) - 1);
    const int64_t *rows = G->dg.dg_row_ptr;
    const int32_t *cols = G->dg.dg_adj_list;
    std::vector<uint64_t> sums(64, 0);
    hipStream_t st = stream();
    H_ERR(hipMemcpyAsync(G->ms_src, sources, sizeof(int32_t) * nsrc,
                         hipMemcpyHostToDevice, st));
    H_ERR(hipEventRecord(G->ev0, st));
    fill_async<uint64_t>(visit, 0, 2 * (n + 1));  // visit and front
    hipLaunchKernelGGL(k_msbfs_seed, dim3(1), dim3(64), 0, st, visit, front,
                       (const int32_t *)G->ms_src, nsrc);
    double front_deg = 0;
    ms_seeds_t seeds{};
    seeds.nsrc = nsrc;
    for (int s = 0; s < nsrc; ++s) {
      const int64_t ds = G->hg.row_ptr[sources[s] + 1] - G->hg.row_ptr[sources[s]];
      seeds.src[s] = sources[s];
      seeds.pre[s + 1] = seeds.pre[s] + ds;
      front_deg += (double)ds;
    }
    G->records.clear();
    int level = 0;
    const int grid = grid_for((n + 63) / 64 * 64);
    while (level < o.max_levels) {
      ++level;
      // direction: forced ("Push-..." / "Pull-..."), else pull once the
      // frontier's edges exceed 1/32 of the graph
      const bool pull = o.force.has_direction ? (o.force.d == Pull)
                                              : front_deg > (double)m / 32.0;
      fill_async<unsigned long long>((unsigned long long *)G->ms_cnt, 0ull,
                                     sizeof(ms_counters_t) / 8 * MS_SLOTS);
      if (pull) {
        // bottom-up over the in-edges (the reverse CSR; the same arrays
        // for an undirected graph)
        hipLaunchKernelGGL(k_msbfs_pull, dim3(grid), dim3(BLOCK), 0, st,
                           (const int64_t *)G->dg.dgr_row_ptr,
                           (const int32_t *)G->dg.dgr_adj_list, rows, n, all,
                           visit, (const uint64_t *)front, next, G->ms_cnt);
      } else if (level == 1) {
        // the sources' edges spread over the grid (they are often hubs)
        fill_async<uint64_t>(next, 0, n + 1);
        hipLaunchKernelGGL(k_msbfs_push_seeds, dim3(grid_for((int64_t)front_deg)),
                           dim3(BLOCK), 0, st, rows, cols, (const uint64_t *)visit,
                           (const uint64_t *)front, next, seeds);
        hipLaunchKernelGGL(k_msbfs_sweep, dim3(grid), dim3(BLOCK), 0, st, rows,
                           n, visit, next, G->ms_cnt);
      } else {
        fill_async<uint64_t>(next, 0, n + 1);
        hipLaunchKernelGGL(k_msbfs_push, dim3(grid), dim3(BLOCK), 0, st, rows,
                           cols, n, (const uint64_t *)visit,
                           (const uint64_t *)front, next);
        hipLaunchKernelGGL(k_msbfs_sweep, dim3(grid), dim3(BLOCK), 0, st, rows,
                           n, visit, next, G->ms_cnt);
      }
      H_ERR(hipGetLastError());
      H_ERR(hipMemcpyAsync(G->ms_hcnt, G->ms_cnt, sizeof(ms_counters_t) * MS_SLOTS,
                           hipMemcpyDeviceToHost, st));
      sync();
      ms_counters_t tot{};
      for (int k = 0; k < MS_SLOTS; ++k) {  // sum the partial slots
        const ms_counters_t &p = G->ms_hcnt[k];
        tot.newly += p.newly;
        tot.newly_deg += p.newly_deg;
        for (int s = 0; s < 64; ++s) tot.per_src[s] += p.per_src[s];
      }
      gsw_level_record rec{};
      rec.level = level;
      rec.dir = pull ? 1 : 0;
      rec.fmt = pull ? 1 : 0;
      rec.frontier = (int64_t)tot.newly;
      rec.edges = (int64_t)tot.newly_deg;
      G->records.push_back(rec);
      if (tot.newly == 0) break;
      for (int s = 0; s < 64; ++s) sums[s] += (uint64_t)level * tot.per_src[s];
      front_deg = (double)tot.newly_deg;
      std::swap(front, next);
    }
    H_ERR(hipEventRecord(G->ev1, st));
    const double ms = elapsed(G);
    if (mask_out)
      H_ERR(hipMemcpy(mask_out, visit, sizeof(uint64_t) * n,
                      hipMemcpyDeviceToHost));
    if (level_sums_out)
      for (int s = 0; s < 64; ++s) level_sums_out[s] = sums[s];
    if (stats) {
      stats->device_ms = ms;
      stats->levels = level;
      stats->n_records = (int32_t)G->records.size();
      stats->reached = 0;
    }
  });
}

// ---- learned selector -------------------------------------------------------
gsw_status gsw_selector_load_app(int32_t app, const float *w, int32_t nparam) {
  return guarded([&] {
    GSW_REQUIRE(app >= 0 && app < GSW_APPS, GSW_E_INVALID, "unknown application");
    GSW_REQUIRE(w && nparam == gsw_mlp::NPARAM, GSW_E_INVALID,
                "expected 6942 weights");
    require_gpu();
    for (int32_t i = 0; i < nparam; ++i)
      GSW_REQUIRE(std::isfinite(w[i]), GSW_E_INVALID, "non-finite weight");
    if (!g_mlp_dev[app])
      H_ERR(hipMalloc((void **)&g_mlp_dev[app], sizeof(float) * gsw_mlp::NPARAM));
    H_ERR(hipMemcpy(g_mlp_dev[app], w, sizeof(float) * gsw_mlp::NPARAM,
                    hipMemcpyHostToDevice));
    g_mlp_host[app].assign(w, w + nparam);
  });
}

gsw_status gsw_selector_load_tree(int32_t app, int32_t nodes, const int32_t *feat,
                                  const float *thr, const int32_t *left,
                                  const int32_t *right, const int32_t *cls) {
  return guarded([&] {
    GSW_REQUIRE(app >= 0 && app < GSW_APPS, GSW_E_INVALID, "unknown application");
    GSW_REQUIRE(nodes >= 1 && nodes <= 4096 && feat && thr && left && right && cls,
                GSW_E_INVALID, "invalid tree");
    for (int32_t i = 0; i < nodes; ++i) {
      if (feat[i] >= 0) {
        // children after their parent: every path ends at a leaf
        GSW_REQUIRE(feat[i] < gsw_mlp::NIN && std::isfinite(thr[i]) && left[i] > i &&
                        left[i] < nodes && right[i] > i && right[i] < nodes,
                    GSW_E_INVALID, "invalid tree node " + std::to_string(i));
      } else {
        GSW_REQUIRE(cls[i] >= 0 && cls[i] < gsw_mlp::NOUT, GSW_E_INVALID,
                    "invalid leaf class at node " + std::to_string(i));
      }
    }
    gsw_tree::table_t &t = g_tree[app];
    t.feat.assign(feat, feat + nodes);
    t.thr.assign(thr, thr + nodes);
    t.left.assign(left, left + nodes);
    t.right.assign(right, right + nodes);
    t.cls.assign(cls, cls + nodes);
  });
}

gsw_status gsw_last_class(const gsw_graph *G, int32_t *cls) {
  return guarded([&] {
    GSW_REQUIRE(G && cls, GSW_E_INVALID, "null argument");
    *cls = G->last_class;
  });
}

gsw_status gsw_selector_load(const float *w, int32_t nparam) {
  return gsw_selector_load_app(GSW_APP_BFS, w, nparam);
}

gsw_status gsw_mlp_features(const double *counts, float *x) {
  return guarded([&] {
    GSW_REQUIRE(counts && x, GSW_E_INVALID, "null argument");
    gsw_mlp::counts_t c;
    c.n = (float)counts[0];
    c.m = (float)counts[1];
    c.avg_deg = (float)counts[2];
    c.std_deg = (float)counts[3];
    c.gini = (float)counts[4];
    c.level = (float)counts[5];
    c.active = (float)counts[6];
    c.inactive = (float)counts[7];
    c.push_wl = (float)counts[8];
    c.pull_wl = (float)counts[9];
    c.max_deg_active = (float)counts[10];
    gsw_mlp::features(c, x);
  });
}

}  // extern "C"

namespace {
// device buffers of one offline MLP call, freed on every exit path
struct dev_bufs_t {
  std::vector<void *> p;
  template <typename T> T *get(size_t n) {
    void *q = nullptr;
    H_ERR(hipMalloc(&q, sizeof(T) * (n ? n : 1)));
    p.push_back(q);
    return (T *)q;
  }
  template <typename T> T *put(const T *h, size_t n) {
    T *d = get<T>(n);
    if (n) H_ERR(hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice));
    return d;
  }
  ~dev_bufs_t() {
    for (void *q : p) hipFree(q);
  }
};

int batch_grid(int64_t B) {
  return (int)MAX((int64_t)1, MIN((int64_t)1024, (B + 3) / 4));
}

// One gradient evaluation over `B` samples (idx == nullptr: off .. off+B)
// into d_grad; returns nothing (loss written to d_loss).
void mlp_grad_device(const float *w, const float *x, const float *t,
                     const float *mask, const int32_t *idx, int64_t off,
                     int64_t B, float scale, float *part, float *part_loss,
                     int nwave_cap, float *d_grad, float *d_loss) {
  constexpr int PER = 8;
  const int64_t nw = MAX((int64_t)1, (B + PER - 1) / PER);
  const int grid = (int)((nw + 3) / 4);
  GSW_REQUIRE(grid * 4 <= nwave_cap, GSW_E_INTERNAL, "gradient scratch too small");
  hipStream_t st = gsw::stream();
  hipLaunchKernelGGL(gsw::k_mlp_grad_partial, dim3(grid), dim3(gsw::BLOCK),
                     gsw::MLP_GRAD_LDS, st, w, x, t, mask, idx, off, B, PER,
                     scale, part, part_loss);
  hipLaunchKernelGGL(gsw::k_mlp_reduce,
                     dim3((gsw_mlp::NPARAM + gsw::BLOCK - 1) / gsw::BLOCK),
                     dim3(gsw::BLOCK), 0, st, part, part_loss, grid * 4, d_grad,
                     d_loss);
  H_ERR(hipGetLastError());
}
}  // namespace

extern "C" {

gsw_status gsw_mlp_forward(const float *w, const float *x, int64_t B, float *y) {
  return guarded([&] {
    GSW_REQUIRE(w && x && y && B >= 0, GSW_E_INVALID, "invalid arguments");
    require_gpu();
    if (B == 0) return;
    dev_bufs_t d;
    const float *dw = d.put(w, gsw_mlp::NPARAM);
    const float *dx = d.put(x, (size_t)B * gsw_mlp::NIN);
    float *dy = d.get<float>((size_t)B * gsw_mlp::NOUT);
    hipLaunchKernelGGL(gsw::k_mlp_forward_batch, dim3(batch_grid(B)),
                       dim3(gsw::BLOCK), 0, gsw::stream(), dw, dx, B, dy);
    H_ERR(hipGetLastError());
    H_ERR(hipMemcpy(y, dy, sizeof(float) * B * gsw_mlp::NOUT,
                    hipMemcpyDeviceToHost));
  });
}

gsw_status gsw_mlp_select(const float *w, const float *x, int64_t B,
                          uint32_t mask, int32_t *cls) {
  return guarded([&] {
    GSW_REQUIRE(w && x && cls && B >= 0, GSW_E_INVALID, "invalid arguments");
    require_gpu();
    if (B == 0) return;
    dev_bufs_t d;
    const float *dw = d.put(w, gsw_mlp::NPARAM);
    const float *dx = d.put(x, (size_t)B * gsw_mlp::NIN);
    int32_t *dc = d.get<int32_t>((size_t)B);
    hipLaunchKernelGGL(gsw::k_mlp_select_batch, dim3(batch_grid(B)),
                       dim3(gsw::BLOCK), 0, gsw::stream(), dw, dx, B, mask, dc);
    H_ERR(hipGetLastError());
    H_ERR(hipMemcpy(cls, dc, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
  });
}

gsw_status gsw_mlp_grad(const float *w, const float *x, const float *t,
                        const float *mask, int64_t B, float *loss,
                        float *grad) {
  return guarded([&] {
    GSW_REQUIRE(w && x && t && mask && grad && B >= 1, GSW_E_INVALID,
                "invalid arguments");
    require_gpu();
    double msum = 0;
    for (int64_t i = 0; i < B * gsw_mlp::NOUT; ++i) msum += mask[i];
    const float scale = msum > 0 ? (float)(1.0 / msum) : 0.f;
    dev_bufs_t d;
    const float *dw = d.put(w, gsw_mlp::NPARAM);
    const float *dx = d.put(x, (size_t)B * gsw_mlp::NIN);
    const float *dt = d.put(t, (size_t)B * gsw_mlp::NOUT);
    const float *dm = d.put(mask, (size_t)B * gsw_mlp::NOUT);
    const int cap = (int)(((B + 7) / 8 + 3) / 4 * 4);
    float *part = d.get<float>((size_t)cap * gsw_mlp::NPARAM);
    float *pl = d.get<float>((size_t)cap);
    float *dg = d.get<float>(gsw_mlp::NPARAM);
    float *dl = d.get<float>(1);
    mlp_grad_device(dw, dx, dt, dm, nullptr, 0, B, scale, part, pl, cap, dg, dl);
    H_ERR(hipMemcpy(grad, dg, sizeof(float) * gsw_mlp::NPARAM,
                    hipMemcpyDeviceToHost));
    float l = 0;
    H_ERR(hipMemcpy(&l, dl, sizeof(float), hipMemcpyDeviceToHost));
    if (loss) *loss = l * scale;
  });
}

gsw_status gsw_mlp_train(const float *x, const float *t, const float *mask,
                         int64_t B, int32_t epochs, int32_t batch, float lr,
                         uint64_t seed, float *w_inout, float *loss_curve) {
  return guarded([&] {
    GSW_REQUIRE(x && t && mask && w_inout && B >= 1 && epochs >= 0 &&
                    batch >= 1 && lr > 0,
                GSW_E_INVALID, "invalid arguments");
    require_gpu();
    using namespace gsw_mlp;
    dev_bufs_t d;
    float *dw = d.put(w_inout, NPARAM);
    const float *dx = d.put(x, (size_t)B * NIN);
    const float *dt = d.put(t, (size_t)B * NOUT);
    const float *dm = d.put(mask, (size_t)B * NOUT);
    float *m1 = d.get<float>(NPARAM);
    float *m2 = d.get<float>(NPARAM);
    gsw::fill_async<float>(m1, 0.f, NPARAM);
    gsw::fill_async<float>(m2, 0.f, NPARAM);
    const int64_t bsz_max = MIN((int64_t)batch, B);
    const int cap = (int)(((bsz_max + 7) / 8 + 3) / 4 * 4);
    float *part = d.get<float>((size_t)cap * NPARAM);
    float *pl = d.get<float>((size_t)cap);
    float *dg = d.get<float>(NPARAM);
    const int64_t steps = (B + batch - 1) / batch;
    float *dl = d.get<float>((size_t)steps);
    int32_t *didx = d.get<int32_t>((size_t)B);
    std::vector<int32_t> perm((size_t)B);
    std::vector<float> hl((size_t)steps);
    // per-sample mask sums for the minibatch scale
    std::vector<double> ms((size_t)B, 0.0);
    double mtot = 0;
    for (int64_t i = 0; i < B; ++i) {
      for (int c = 0; c < NOUT; ++c) ms[i] += mask[i * NOUT + c];
      mtot += ms[i];
    }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int64_t step = 0;
    const uint64_t key = gsw_host::stream_key(seed, 11);
    for (int e = 0; e < epochs; ++e) {
      for (int64_t i = 0; i < B; ++i) perm[i] = (int32_t)i;
      for (int64_t i = B - 1; i > 0; --i) {  // Fisher-Yates, counter-based draws
        const uint64_t r = gsw_host::draw(key, (uint64_t)e * (uint64_t)B + (uint64_t)i);
        std::swap(perm[i], perm[gsw_host::below(r, (uint64_t)i + 1)]);
      }
      H_ERR(hipMemcpy(didx, perm.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice));
      for (int64_t s = 0; s < steps; ++s) {
        const int64_t off = s * batch, bsz = MIN((int64_t)batch, B - off);
        double msum = 0;
        for (int64_t i = 0; i < bsz; ++i) msum += ms[perm[off + i]];
        const float scale = msum > 0 ? (float)(1.0 / msum) : 0.f;
        mlp_grad_device(dw, dx, dt, dm, didx, off, bsz, scale, part, pl, cap, dg,
                        dl + s);
        ++step;
        const double bc1 = 1.0 - std::pow((double)b1, (double)step);
        const double bc2 = 1.0 - std::pow((double)b2, (double)step);
        const float lr_t = (float)(lr * std::sqrt(bc2) / bc1);
        hipLaunchKernelGGL(gsw::k_adam, dim3((NPARAM + gsw::BLOCK - 1) / gsw::BLOCK),
                           dim3(gsw::BLOCK), 0, gsw::stream(), dw, dg, m1, m2,
                           NPARAM, lr_t, b1, b2, eps);
        H_ERR(hipGetLastError());
      }
      if (loss_curve) {
        H_ERR(hipMemcpy(hl.data(), dl, sizeof(float) * steps, hipMemcpyDeviceToHost));
        double s = 0;
        for (float v : hl) s += v;
        loss_curve[e] = mtot > 0 ? (float)(s / mtot) : 0.f;
      }
    }
    H_ERR(hipMemcpy(w_inout, dw, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  });
}

// ---- exhaustive per-level timing (selector training labels) ----------------
gsw_status gsw_explore_bfs(gsw_graph *G, int32_t root, int32_t max_levels,
                           int32_t stride, int32_t reps, float *feat_out,
                           float *ms_out, int32_t *rule_out, int32_t *level_out,
                           int32_t *n_out) {
  return guarded([&] {
    using namespace gsw_mlp;
    GSW_REQUIRE(G && n_out && max_levels >= 1 && stride >= 1 && reps >= 1,
                GSW_E_INVALID, "invalid arguments");
    GSW_REQUIRE(root >= 0 && root < G->hg.n, GSW_E_INVALID, "root out of range");
    auto &c = bfs_ctx(G);
    arch_once(c);
    const int64_t n = G->hg.n;
    dev_bufs_t d;
    int32_t *snap = d.get<int32_t>((size_t)n);
    hipStream_t st = gsw::stream();
    init_bfs_levels(G, c.f.data.dg_wa, root, nullptr, 1);
    int written = 0;
    for (int lvl = 0; lvl < (1 << 20) && written < max_levels; ++lvl) {
      H_ERR(hipMemcpyAsync(snap, c.f.data.dg_wa, sizeof(int32_t) * n,
                           hipMemcpyDeviceToDevice, st));
      // the level as the rule selector sees it
      bfs_level_setup(G, c, root, -1);
      GU g = G->dg;
      g.level = lvl;
      c.en.inspector.inspect(c.as, g, c.f, c.en.stats, c.en.fets, c.en.conf);
      const long long *info = c.en.stats.h_runtime_info;
      // the engine stops here too: no frontier, or nothing left to reach
      if (info[0] == 0 || info[1] == 0) break;
      const bool probe = (lvl % stride) == 0;
      if (!probe) {
        c.en.selector.select(c.en.stats, c.en.fets, c.en.conf);
        c.en.executor.filter(c.as, g, c.f, c.en.stats, c.en.fets, c.en.conf);
        g.update_level();
        c.en.executor.expand(c.as, g, c.f, c.en.stats, c.en.fets, c.en.conf);
        gsw::sync();
        continue;
      }
      counts_t k;
      k.n = (float)n;
      k.m = (float)G->hg.m;
      k.avg_deg = (float)G->attr.avg_deg;
      k.std_deg = (float)G->attr.std_deg;
      k.gini = (float)G->attr.GI;
      k.level = (float)lvl;
      k.active = (float)info[0];
      k.inactive = (float)info[1];
      k.push_wl = (float)info[2];
      k.pull_wl = (float)info[3];
      k.max_deg_active = (float)info[4];
      if (feat_out) features(k, feat_out + (size_t)written * NIN);
      c.en.selector.select(c.en.stats, c.en.fets, c.en.conf);
      const config_t &rc = c.en.conf;
      if (rule_out)
        rule_out[written] = class_of(rc.conf_dir, rc.conf_asfmt, rc.conf_sorted,
                                     rc.conf_fuse_inspect, rc.conf_lb);
      if (level_out) level_out[written] = lvl;
      for (int cls = 0; cls < NOUT; ++cls) {
        double best = 1e300;
        for (int r = 0; r < reps; ++r) {
          H_ERR(hipMemcpyAsync(c.f.data.dg_wa, snap, sizeof(int32_t) * n,
                               hipMemcpyDeviceToDevice, st));
          // one probe per restore (a macro MIN would run the probe twice)
          best = std::min(best, bfs_probe_level(G, c, root, cls, lvl));
        }
        if (ms_out) ms_out[(size_t)written * NOUT + cls] = (float)best;
      }
      // every class leaves the same levels array: the state of level lvl+1
      ++written;
    }
    *n_out = written;
  });
}

gsw_status gsw_explore_sssp(gsw_graph *G, int32_t root, int32_t max_levels,
                            int32_t stride, int32_t reps, float *feat_out,
                            float *ms_out, int32_t *rule_out, int32_t *level_out,
                            int32_t *n_out) {
  return guarded([&] {
    using namespace gsw_mlp;
    GSW_REQUIRE(G && n_out && max_levels >= 1 && stride >= 1 && reps >= 1,
                GSW_E_INVALID, "invalid arguments");
    GSW_REQUIRE(G->has_w, GSW_E_INVALID, "SSSP needs a weighted graph");
    GSW_REQUIRE(root >= 0 && root < G->hg.n, GSW_E_INVALID, "root out of range");
    auto &c = sssp_ctx(G);
    arch_once(c);
    const int64_t n = G->hg.n;
    dev_bufs_t d;
    uint32_t *snap_wa = d.get<uint32_t>((size_t)n);
    uint32_t *snap_ra = d.get<uint32_t>((size_t)n);
    hipStream_t st = gsw::stream();
    // the state the level loop starts from
    sssp_prepare(G, c, run_opts_t{}, root, false);
    feature_t fets = c.en.fets;
    window_t win = c.f.data.window;
    hipLaunchKernelGGL(k_init_sssp, dim3(gsw::grid_for(n)), dim3(gsw::BLOCK), 0, st,
                       c.f.data.dg_wa, c.f.data.dg_ra, n, root);
    H_ERR(hipGetLastError());
    auto save = [&] {
      H_ERR(hipMemcpyAsync(snap_wa, c.f.data.dg_wa, sizeof(uint32_t) * n,
                           hipMemcpyDeviceToDevice, st));
      H_ERR(hipMemcpyAsync(snap_ra, c.f.data.dg_ra, sizeof(uint32_t) * n,
                           hipMemcpyDeviceToDevice, st));
    };
    auto restore = [&] {
      H_ERR(hipMemcpyAsync(c.f.data.dg_wa, snap_wa, sizeof(uint32_t) * n,
                           hipMemcpyDeviceToDevice, st));
      H_ERR(hipMemcpyAsync(c.f.data.dg_ra, snap_ra, sizeof(uint32_t) * n,
                           hipMemcpyDeviceToDevice, st));
    };
    int written = 0;
    // lvl counts expanded levels, as the level loop's graph level does
    // (an iteration that only advances the window expands nothing)
    int lvl = 0;
    for (int it = 0; it < (1 << 20) && written < max_levels; ++it) {
      save();
      // the level as the rule selector runs it
      sssp_level_setup(G, c, root, -1, lvl, fets, win);
      GW g = G->dgw;
      g.level = lvl;
      c.en.inspector.inspect(c.as, g, c.f, c.en.stats, c.en.fets, c.en.conf);
      if (c.as.finish(g, c.f, c.en.conf)) break;  // no pending update left
      if (lvl > 0 && c.en.fets.active_vertex <= 0) {
        // the window advanced over a drained frontier: the next level
        // inspects with the new window
        fets = c.en.fets;
        win = c.f.data.window;
        continue;
      }
      if ((lvl % stride) == 0) {
        const feature_t lf = c.en.fets;
        counts_t k;
        k.n = (float)n;
        k.m = (float)G->hg.m;
        k.avg_deg = (float)G->attr.avg_deg;
        k.std_deg = (float)G->attr.std_deg;
        k.gini = (float)G->attr.GI;
        k.level = (float)lvl;
        k.active = (float)lf.active_vertex;
        k.inactive = (float)lf.unactive_vertex;
        k.push_wl = (float)lf.push_workload;
        k.pull_wl = (float)lf.pull_workload;
        k.max_deg_active = (float)lf.cur_max_deg_active;
        if (feat_out) features(k, feat_out + (size_t)written * NIN);
        c.en.selector.select(c.en.stats, c.en.fets, c.en.conf);
        const config_t &rc = c.en.conf;
        if (rule_out)
          rule_out[written] = class_of(rc.conf_dir, rc.conf_asfmt, rc.conf_sorted,
                                       rc.conf_fuse_inspect, rc.conf_lb);
        if (level_out) level_out[written] = lvl;
        for (int cls = 0; cls < NOUT; ++cls) {
          double best = 1e300;
          for (int r = 0; r < reps; ++r) {
            restore();
            best = std::min(best, sssp_probe_level(G, c, root, cls, lvl, fets, win));
          }
          if (ms_out) ms_out[(size_t)written * NOUT + cls] = (float)best;
        }
        ++written;
        // back to the level's start for the rule's run of it
        restore();
        sssp_level_setup(G, c, root, -1, lvl, fets, win);
        c.en.inspector.inspect(c.as, g, c.f, c.en.stats, c.en.fets, c.en.conf);
        c.as.finish(g, c.f, c.en.conf);
      }
      c.en.selector.select(c.en.stats, c.en.fets, c.en.conf);
      c.en.executor.filter(c.as, g, c.f, c.en.stats, c.en.fets, c.en.conf);
      g.update_level();
      c.en.executor.expand(c.as, g, c.f, c.en.stats, c.en.fets, c.en.conf);
      gsw::sync();
      fets = c.en.fets;
      win = c.f.data.window;
      ++lvl;
    }
    *n_out = written;
  });
}

gsw_status gsw_tree_eval(const char *name, const double *fv, int32_t nfv,
                         int32_t *cls) {
  return guarded([&] {
    GSW_REQUIRE(name && fv && cls, GSW_E_INVALID, "invalid arguments");
    using namespace gsw_model;
    struct {
      const char *n;
      const tree_t *t;
    } table[] = {
        {"select_dir_idem", &k_select_dir_idem},
        {"select_dir_asso", &k_select_dir_asso},
        {"select_dir_asso_backup", &k_select_dir_asso_backup},
        {"select_fmt", &k_select_fmt},
        {"select_fusion", &k_select_fusion},
        {"select_lb_idem", &k_select_lb_idem},
        {"select_lb_asso_forward", &k_select_lb_asso_forward},
        {"select_lb_asso_backward", &k_select_lb_asso_backward},
        {"select_lb_asso_forward_backup", &k_select_lb_asso_forward_backup},
        {"select_lb_asso_backward_backup", &k_select_lb_asso_backward_backup},
        {"select_stepping", &k_select_stepping},
    };
    for (auto &e : table) {
      if (!std::strcmp(e.n, name)) {
        GSW_REQUIRE(nfv >= e.t->n_features, GSW_E_INVALID,
                    "feature vector too short for this tree");
        *cls = eval_tree(*e.t, fv);
        return;
      }
    }
    throw gsw::error(GSW_E_INVALID, std::string("unknown tree ") + name);
  });
}

// ---- primitives --------------------------------------------------------------
namespace {
__global__ void k_wave_scan64(const int *in, int *out) {
  out[threadIdx.x] = gsw::wave_inclusive_scan(in[threadIdx.x]);
}
}  // namespace

gsw_status gsw_wave_scan64(const int32_t *in, int32_t *out) {
  return guarded([&] {
    GSW_REQUIRE(in && out, GSW_E_INVALID, "invalid arguments");
    require_gpu();
    int *d = nullptr;
    H_ERR(hipMalloc((void **)&d, sizeof(int) * 128));
    H_ERR(hipMemcpy(d, in, sizeof(int) * 64, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_wave_scan64, dim3(1), dim3(64), 0, gsw::stream(), d,
                       d + 64);
    H_ERR(hipGetLastError());
    H_ERR(hipMemcpy(out, d + 64, sizeof(int) * 64, hipMemcpyDeviceToHost));
    hipFree(d);
  });
}

namespace {
__global__ void __launch_bounds__(gsw::BLOCK)
    k_stream_copy(const int4 *__restrict__ src, int4 *__restrict__ dst,
                  int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

gsw_status gsw_stream_copy(int64_t bytes, int32_t reps, double *gbps) {
  return guarded([&] {
    GSW_REQUIRE(gbps && bytes >= 16 && reps >= 1, GSW_E_INVALID,
                "invalid arguments");
    require_gpu();
    const int64_t n16 = bytes / 16;
    int4 *a = nullptr, *b = nullptr;
    H_ERR(hipMalloc((void **)&a, n16 * 16));
    H_ERR(hipMalloc((void **)&b, n16 * 16));
    H_ERR(hipMemset(a, 1, n16 * 16));
    hipEvent_t e0, e1;
    H_ERR(hipEventCreate(&e0));
    H_ERR(hipEventCreate(&e1));
    hipStream_t st = gsw::stream();
    const int grid = 256 * 8 * 4;
    double best = 0;
    for (int r = 0; r < reps; ++r) {
      H_ERR(hipEventRecord(e0, st));
      hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(gsw::BLOCK), 0, st, a,
                         b, n16);
      H_ERR(hipEventRecord(e1, st));
      H_ERR(hipEventSynchronize(e1));
      float ms = 0;
      H_ERR(hipEventElapsedTime(&ms, e0, e1));
      const double g = 2.0 * (double)n16 * 16 / (ms * 1e-3) / 1e9;
      best = g > best ? g : best;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    *gbps = best;
  });
}

gsw_status gsw_scan_exclusive_i64(const int64_t *in, int64_t n, int64_t *out) {
  return guarded([&] {
    GSW_REQUIRE(in && out && n >= 0, GSW_E_INVALID, "invalid arguments");
    require_gpu();
    int64_t *d = nullptr, *bs = nullptr;
    const int64_t nb = (n + gsw::SCAN_TILE - 1) / gsw::SCAN_TILE;
    H_ERR(hipMalloc((void **)&d, sizeof(int64_t) * (n + 1)));
    H_ERR(hipMalloc((void **)&bs, sizeof(int64_t) * (nb + 2)));
    if (n) H_ERR(hipMemcpy(d, in, sizeof(int64_t) * n, hipMemcpyHostToDevice));
    gsw::scan_exclusive_inplace(d, n, bs);
    H_ERR(hipMemcpy(out, d, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    hipFree(d);
    hipFree(bs);
  });
}

}  // extern "C"

// ---- allocation-free connected components -------------------------------
namespace {

// Scratch of gsw_cc_noalloc: the sampled vertex ids, then the chosen root.
// Every GPU holds its own copy for the life of the process. In host code
// the name is only a stand-in for the symbol: kernels get the device address
// that hipGetSymbolAddress returns for the current device.
constexpr int NOALLOC_SCRATCH_INTS = gsw::CC_MAX_SAMPLES + 64;
static __device__ __align__(256) int32_t scratch[NOALLOC_SCRATCH_INTS];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
int32_t *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// address of `scratch` on the current device, looked up once per device
int32_t *noalloc_scratch() {
  int dev = 0;
  H_ERR(hipGetDevice(&dev));
  GSW_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, GSW_E_UNSUPPORTED,
              "device id beyond the scratch address table");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *p = nullptr;
    H_ERR(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    GSW_REQUIRE(p != nullptr, GSW_E_HIP, "no device address for the CC scratch");
    g_noalloc_ptr[dev] = (int32_t *)p;
  }
  return g_noalloc_ptr[dev];
}

}  // namespace

extern "C" {

gsw_status gsw_cc_noalloc(const gsw_graph *G, int32_t samples, int32_t *labels,
                          void *stream) {
  return guarded([&] {
    GSW_REQUIRE(G && labels, GSW_E_INVALID, "invalid arguments");
    GSW_REQUIRE(samples >= 1 && samples <= gsw::CC_MAX_SAMPLES, GSW_E_INVALID,
                "samples must be 1.." + std::to_string(gsw::CC_MAX_SAMPLES) +
                    " (the size of the static scratch)");
    hipPointerAttribute_t attr;
    const hipError_t pe = hipPointerGetAttributes(&attr, labels);
    if (pe != hipSuccess) (void)hipGetLastError();  // clear the failed query
    GSW_REQUIRE(pe == hipSuccess &&
                    (attr.type == hipMemoryTypeDevice ||
                     attr.type == hipMemoryTypeManaged),
                GSW_E_INVALID, "labels must be device memory");
    int32_t *s = noalloc_scratch();
    cc_sampled_launch(G, labels, s, samples, (hipStream_t)stream);
  });
}

}  // extern "C"
