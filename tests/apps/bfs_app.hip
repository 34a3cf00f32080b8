// BFS written against the GSWITCH template API the way a GSWITCH
// application is written (application/bfs.cu:1-153 is the shape): global
// actors, a Functor<VC>, parse_cmd, build_graph from a file, the host level
// loop inspect -> finish -> select -> filter -> update_level -> expand ->
// finish, and a CPU check behind -V. Only the include path and the HIP
// runtime differ from a reference application. For the tests,
// GSW_APP_OUT=<file> also writes the root and the levels (int32) there.
#include <cstdio>
#include <cstdlib>
#include <queue>
#include <vector>

#include "gswitch.h"

using G = device_graph_t<CSR, Empty>;

inspector_t inspector;
selector_t selector;
executor_t executor;
feature_t fets;
config_t conf;
stat_t stats;

struct Levels : Functor<VC, int, Empty, Empty> {
  __device__ Status filter(int v, G g) {
    const int l = *wa_of(v);
    if (l == g.get_level()) return Active;
    return l < 0 ? Inactive : Fixed;
  }
  __device__ int emit(int v, Empty *e, G g) { return g.get_level(); }
  __device__ bool cond(int v, int msg, G g) { return *wa_of(v) < 0; }
  __device__ bool comp(int *d, int msg, G g) {
    *d = msg;
    return true;
  }
  __device__ bool compAtomic(int *d, int msg, G g) {
    *d = msg;
    return true;
  }
};

// plain queue BFS on the host graph
std::vector<int> cpu_levels(host_graph_t<CSR, Empty> hg, int root) {
  std::vector<int> lv(hg.nvertexs, -1);
  std::queue<int> q;
  lv[root] = 0;
  q.push(root);
  while (!q.empty()) {
    const int v = q.front();
    q.pop();
    for (int64_t j = hg.start_pos[v]; j < hg.start_pos[v + 1]; ++j) {
      const int u = hg.adj_list[j];
      if (lv[u] < 0) {
        lv[u] = lv[v] + 1;
        q.push(u);
      }
    }
  }
  return lv;
}

template <typename Gd, typename F> double traverse(Gd g, F f, int root) {
  active_set_t as = build_active_set(g.nvertexs, conf);
  as.init(root);
  const double t0 = mwtime();
  for (int level = 0;; ++level) {
    inspector.inspect(as, g, f, stats, fets, conf);
    if (as.finish(g, f, conf)) break;
    selector.select(stats, fets, conf);
    executor.filter(as, g, f, stats, fets, conf);
    g.update_level();
    executor.expand(as, g, f, stats, fets, conf);
    if (as.finish(g, f, conf)) break;
  }
  return mwtime() - t0;
}

int main(int argc, char *argv[]) {
  parse_cmd(argc, argv, "BFS");
  fets.centric = VC;
  fets.pattern = Idem;
  fets.fromall = false;
  fets.toall = false;
  conf.conf_pruning = true;

  auto g = build_graph<VC>(cmd_opt.path, fets, cmd_opt.with_header,
                           cmd_opt.with_weight, cmd_opt.directed);
  if (g.hg.nedges == 0) return 1;
  Levels f;
  f.data.build(g.hg.nvertexs);
  int root = cmd_opt.src;
  if (root < 0) root = g.hg.random_root();
  LOG(" -- root %d\n", root);
  fets.use_root = root;
  f.data.init_wa([root](int i) { return i == root ? 0 : -1; });
  init_conf(stats, fets, conf, g, f);

  const double ms = traverse(g.dg, f, root);
  f.data.sync_wa();
  if (cmd_opt.validation) {
    const std::vector<int> ref = cpu_levels(g.hg, root);
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
