// Connected components written against the GSWITCH template API in a
// GSWITCH application's shape (application/cc.cu:1-140): an edge-centric
// Functor<EC> on the COO graph from build_graph<EC>, hooking levels with
// p_jump after every expand, and a CPU check behind -V that builds its own
// CSR from the graph's edge list (host_graph_t::build(el, quiet)).
// GSW_APP_OUT=<file> also writes the labels (int32) there.
#include <cstdio>
#include <cstdlib>
#include <queue>
#include <vector>

#include "gswitch.h"

inspector_t inspector;
selector_t selector;
executor_t executor;
feature_t fets;
config_t conf;
stat_t stats;

struct Components : Functor<EC, int, Empty, Empty> {
  __device__ Status filter(int v, int u, Empty *e) {
    return *wa_of(v) == *wa_of(u) ? Fixed : Active;
  }
  __device__ void update(int v, int u, Empty *e) {
    int *rv = wa_of(*wa_of(v));
    int *ru = wa_of(*wa_of(u));
    if (*rv < *ru)
      atomicMin(ru, *rv);
    else if (*ru < *rv)
      atomicMin(rv, *ru);
  }
};

// label every vertex with the smallest id of its component
std::vector<int> cpu_labels(const edgelist_t<Empty> &el) {
  host_graph_t<CSR, Empty> hg;
  hg.build(el, true);
  std::vector<int> id(hg.nvertexs);
  for (int64_t i = 0; i < hg.nvertexs; ++i) id[i] = (int)i;
  std::queue<int> q;
  for (int64_t i = 0; i < hg.nvertexs; ++i) {
    if (id[i] != i) continue;
    q.push((int)i);
    while (!q.empty()) {
      const int v = q.front();
      q.pop();
      for (int64_t j = hg.start_pos[v]; j < hg.start_pos[v + 1]; ++j) {
        const int u = hg.adj_list[j];
        if (id[u] == u && u != i) {
          id[u] = (int)i;
          q.push(u);
        }
      }
    }
  }
  return id;
}

template <typename Gd, typename F> double components(Gd g, F f) {
  active_set_t as = build_active_set(g.nedges, conf);
  as.init(ALL_ACTIVE);
  int *dg_flag = nullptr, *h_flag = nullptr;
  H_ERR(hipMalloc((void **)&dg_flag, sizeof(int)));
  H_ERR(hipHostMalloc((void **)&h_flag, sizeof(int)));
  const double t0 = mwtime();
  for (int level = 0;; ++level) {
    inspector.inspect(as, g, f, stats, fets, conf);
    if (as.finish(g, f, conf)) break;
    selector.select(stats, fets, conf);
    executor.filter(as, g, f, stats, fets, conf);
    g.update_level();
    executor.expand(as, g, f, stats, fets, conf);
    p_jump(f, g.nvertexs, dg_flag, h_flag);
    if (as.finish(g, f, conf)) break;
  }
  return mwtime() - t0;
}

int main(int argc, char *argv[]) {
  parse_cmd(argc, argv, "CC");
  fets.centric = EC;
  fets.pattern = ASSO;

  auto g = build_graph<EC>(cmd_opt.path, fets, cmd_opt.with_header,
                           cmd_opt.with_weight, cmd_opt.directed);
  Components f;
  f.data.build(g.hg.nvertexs);
  f.data.init_wa([](int i) { return i; });
  init_conf(stats, fets, conf, g, f);

  const double ms = components(g.dg, f);
  f.data.sync_wa();
  if (cmd_opt.validation) {
    const std::vector<int> ref = cpu_labels(g.el);
    bool ok = true;
    for (int64_t v = 0; v < g.hg.nvertexs && ok; ++v) ok = ref[v] == f.data.h_wa[v];
    puts(ok ? "passed" : "failed");
  }
  if (const char *out = std::getenv("GSW_APP_OUT")) {
    FILE *fp = std::fopen(out, "wb");
    std::fwrite(f.data.h_wa, sizeof(int), g.hg.nvertexs, fp);
    std::fclose(fp);
  }
  std::printf("%.3f\n", ms);
  return 0;
}
