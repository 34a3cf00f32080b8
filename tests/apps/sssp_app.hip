// SSSP written against the GSWITCH template API in a GSWITCH application's
// shape (application/sssp.cu:1-180): int distances, a delta-stepping window
// (conf.conf_window), an edge list read with edgelist_t::read_mtx, weights
// drawn by gen_weight(64) unless the file has them (-W), build_graph from
// the edge list, the level loop inspect -> select -> filter -> update_level
// -> expand -> finish, and an SPFA check behind -V. GSW_APP_OUT=<file> also
// writes the root and the distances (int32, INT_MAX unreached) there.
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <queue>
#include <vector>

#include "gswitch.h"

using G = device_graph_t<CSR, int>;
const int maxi = std::numeric_limits<int>::max();

inspector_t inspector;
selector_t selector;
executor_t executor;
feature_t fets;
config_t conf;
stat_t stats;

struct Dist : Functor<VC, int, int, int> {
  __device__ Status filter(int v, G g) {
    int *now = wa_of(v);
    int *was = ra_of(v);
    if (*now == *was) return Inactive;
    if (data.window.in(*now)) {
      *was = *now;
      return Active;
    }
    return Inactive;
  }
  __device__ int emit(int v, int *w, G g) { return *wa_of(v) + *w; }
  __device__ bool comp(int *d, int msg, G g) {
    if (*d > msg) {
      *d = msg;
      return data.window.in(msg);
    }
    return false;
  }
  __device__ bool compAtomic(int *d, int msg, G g) {
    const int old = atomicMin(d, msg);
    return old > msg && data.window.in(msg);
  }
  __device__ bool cond(int v, int msg, G g) { return *wa_of(v) > msg; }
  __device__ bool exit(int v, G g) { return *wa_of(v) != *ra_of(v); }
};

// label-correcting queue on the host graph
std::vector<int> cpu_dist(host_graph_t<CSR, int> hg, int root) {
  std::vector<int> d(hg.nvertexs, maxi);
  std::vector<char> inq(hg.nvertexs, 0);
  std::queue<int> q;
  d[root] = 0;
  q.push(root);
  inq[root] = 1;
  while (!q.empty()) {
    const int v = q.front();
    q.pop();
    inq[v] = 0;
    for (int64_t j = hg.start_pos[v]; j < hg.start_pos[v + 1]; ++j) {
      const int u = hg.adj_list[j], nd = d[v] + hg.edge_weights[j];
      if (nd < d[u]) {
        d[u] = nd;
        if (!inq[u]) {
          inq[u] = 1;
          q.push(u);
        }
      }
    }
  }
  return d;
}

template <typename Gd, typename F> double traverse(Gd g, F f, int root) {
  active_set_t as = build_active_set(g.nvertexs, conf);
  as.init(root);
  const double t0 = mwtime();
  for (int level = 0;; ++level) {
    inspector.inspect(as, g, f, stats, fets, conf);
    selector.select(stats, fets, conf);
    executor.filter(as, g, f, stats, fets, conf);
    g.update_level();
    executor.expand(as, g, f, stats, fets, conf);
    if (as.finish(g, f, conf)) break;
  }
  return mwtime() - t0;
}

int main(int argc, char *argv[]) {
  parse_cmd(argc, argv, "SSSP");
  fets.centric = VC;
  fets.pattern = Mono;
  fets.fromall = false;
  fets.toall = true;
  conf.conf_window = true;

  edgelist_t<int> el;
  el.read_mtx(cmd_opt.path, cmd_opt.directed, cmd_opt.with_weight,
              cmd_opt.with_header);
  el.gen_weight(64);
  auto g = build_graph<VC, int>(el, fets);
  Dist f;
  f.data.build(g.hg.nvertexs);
  int root = cmd_opt.src;
  if (root < 0) root = g.hg.random_root();
  LOG(" -- root %d\n", root);
  fets.use_root = root;
  f.data.init_wa([root](int i) { return i == root ? 0 : maxi; });
  f.data.init_ra([root](int i) { return i == root ? 0 : maxi; });
  init_conf(stats, fets, conf, g, f);

  const double ms = traverse(g.dg, f, root);
  f.data.sync_wa();
  if (cmd_opt.validation) {
    const std::vector<int> ref = cpu_dist(g.hg, root);
    bool ok = true;
    for (int64_t v = 0; v < g.hg.nvertexs && ok; ++v) ok = ref[v] == f.data.h_wa[v];
    puts(ok ? "passed" : "failed");
  }
  if (const char *out = std::getenv("GSW_APP_OUT")) {
    FILE *fp = std::fopen(out, "wb");
    std::fwrite(&root, sizeof(int), 1, fp);
    std::fwrite(f.data.h_wa, sizeof(int), g.hg.nvertexs, fp);
    std::fclose(fp);
  }
  std::printf("%.3f\n", ms);
  return 0;
}
