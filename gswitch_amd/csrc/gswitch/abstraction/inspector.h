// inspector_t: the per-level "inspect" actor.
// Follows inspector_t (src/abstraction/inspector.cuh:112-244): skip the pass
// in fused mode and in the first round (features are synthesised from the
// root), otherwise run the inspector kernel, read the six counters back
// (one pinned copy + one stream sync) and turn them into features; declare
// convergence when nothing is active (or, for idempotent algorithms, when
// nothing is left inactive).
#pragma once
#include <chrono>

#include "../data_structures/active_set.h"
#include "../kernels/inspect.h"
#include "features.h"
#include "statistics.h"

inline double gsw_now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch())
      .count();
}

struct inspector_t {
  // Selector inference on the host for a level whose counters are known
  // there: the first round (network or class tree), and every inspected
  // level under the class tree. Same features as the device path.
  void mlp_host_select(const stat_t &stat, const feature_t &fets,
                       config_t &conf, int level) {
    const gsw_mlp::sel_dev_t &mlp = stat.mlp;
    gsw_mlp::counts_t c;
    c.n = mlp.n;
    c.m = mlp.m;
    c.avg_deg = mlp.avg_deg;
    c.std_deg = mlp.std_deg;
    c.gini = mlp.gini;
    c.level = (float)level;
    c.active = (float)fets.active_vertex;
    c.inactive = (float)fets.unactive_vertex;
    c.push_wl = (float)fets.push_workload;
    c.pull_wl = (float)fets.pull_workload;
    c.max_deg_active = (float)fets.cur_max_deg_active;
    float x[gsw_mlp::NIN], y[gsw_mlp::NOUT];
    gsw_mlp::features(c, x);
    if (stat.tree) {
      const int k = stat.tree->eval(x);
      // a leaf class the application cannot run falls back to the rules
      conf.mlp_class = (k >= 0 && k < gsw_mlp::NOUT && ((mlp.mask >> k) & 1u)) ? k : -1;
      return;
    }
    gsw_mlp::forward_host(stat.mlp_host_w, x, y);
    conf.mlp_class = gsw_mlp::argmin_valid(y, mlp.mask);
  }

  template <typename E, typename F>
  void inspect(active_set_t &as, device_graph_t<CSR, E> &g, F &f,
               stat_t &stat, feature_t &fets, config_t &conf) {
    stat.last_filter_time = gsw_now_ms();
    // The window moves only on inspected levels: a fused phase keeps the
    // window its queue was built with, so every in-window update is queued
    // and finish() may absorb the stale marks when the phase drains. (The
    // rule trees' fused phases sit at the window limit anyway, where the
    // reference's per-level adjust changes nothing.)
    if (!conf.conf_fuse_inspect)
      f.data.window.adjust(fets.active_vertex, fets.cur_avg_deg_active,
                           conf.conf_fusion);
    bool need = !conf.conf_fuse_inspect;
    const bool pre = conf.conf_pre_inspected;
    conf.conf_pre_inspected = false;
    init_fets(fets, (double)g.nvertexs, (double)g.nedges);
    if (conf.conf_first_round) {
      // the initial frontier is known exactly (root or everything), also
      // when the first level is fused and features are not synthesised
      as.known_active = fets.use_root < 0 ? g.nvertexs : (long long)fets.first_active;
      as.known_inactive = g.nvertexs - as.known_active;
    }
    if (need && conf.conf_first_round) {
      set_first_fets(as, fets, conf, (double)g.nvertexs, (double)g.nedges);
      if (conf.policy_mlp && stat.learned()) mlp_host_select(stat, fets, conf, g.level);
      need = false;
    }
    if (need && conf.counts_known) {
      // (the filter pass is skipped when the last expand applied it)
      if (!conf.filter_done) {
        hipLaunchKernelGGL((gsw::k_inspect_vc<device_graph_t<CSR, E>, F>),
                           dim3(gsw::grid_for((g.nvertexs + 63) / 64 * 64)),
                           dim3(gsw::BLOCK), 0, gsw::stream(), as, g, f,
                           stat.slots());
        H_ERR(hipGetLastError());
      }
      conf.filter_done = false;
      set_first_fets(as, fets, conf, (double)g.nvertexs, (double)g.nedges);
      need = false;
    }
    if (need && pre) {
      // the previous expand already classified this level (bitmaps written,
      // counters and, under the learned selector, the class in stat)
      stat.collect();
      if (conf.policy_mlp && stat.mlp.w) conf.mlp_class = (int)stat.h_runtime_info[6];
      set_fets(as, stat, fets, conf, (double)g.nvertexs, (double)g.nedges);
      if (conf.policy_mlp && stat.tree) mlp_host_select(stat, fets, conf, g.level);
      need = false;
    }
    if (need) {
      // the partial slots are zero between passes: k_stat_finish clears them
      gsw_mlp::sel_dev_t d;
      // (a run that keeps its first choice needs no inference after it)
      if (conf.policy_mlp && stat.mlp.w && !conf.conf_inherit &&
          !conf.mlp_pr_classes) {
        d = stat.mlp;
        d.level = (float)g.level;
      }
      hipLaunchKernelGGL((gsw::k_inspect_vc<device_graph_t<CSR, E>, F>),
                         dim3(gsw::grid_for((g.nvertexs + 63) / 64 * 64)),
                         dim3(gsw::BLOCK), 0, gsw::stream(), as, g, f,
                         stat.slots());
      hipLaunchKernelGGL(gsw::k_stat_finish, dim3(1), dim3(gsw::BLOCK), 0,
                         gsw::stream(), d, stat.slots(), stat.dg_runtime_info);
      H_ERR(hipGetLastError());
      stat.collect();
      if (d.w) conf.mlp_class = (int)stat.h_runtime_info[6];
      set_fets(as, stat, fets, conf, (double)g.nvertexs, (double)g.nedges);
      if (conf.policy_mlp && stat.tree) mlp_host_select(stat, fets, conf, g.level);
    }
    fets.flatten();
  }

  template <typename E, typename F>
  void inspect(active_set_t &as, device_graph_t<COO, E> &g, F &f,
               stat_t &stat, feature_t &fets, config_t &conf) {
    stat.last_filter_time = gsw_now_ms();
    init_fets(fets, (double)g.nvertexs, (double)g.nedges);
    if (conf.conf_first_round) {
      conf.conf_first_round = false;
      fets.active_vertex = (double)g.nedges;
      fets.active_vertex_ratio = 1;
      as.known_active = g.nedges;
      as.known_inactive = 0;
      return;
    }
    hipLaunchKernelGGL((gsw::k_inspect_ec<device_graph_t<COO, E>, F>),
                       dim3(gsw::grid_for((g.nedges + 63) / 64 * 64)),
                       dim3(gsw::BLOCK), 0, gsw::stream(), as, g, f,
                       stat.slots());
    hipLaunchKernelGGL(gsw::k_stat_finish, dim3(1), dim3(gsw::BLOCK), 0,
                       gsw::stream(), gsw_mlp::sel_dev_t{}, stat.slots(),
                       stat.dg_runtime_info);
    H_ERR(hipGetLastError());
    stat.collect();
    set_fets(as, stat, fets, conf, (double)g.nvertexs, (double)g.nedges);
    fets.flatten();
  }

  void init_fets(feature_t &fets, double n, double m) {
    fets.active_vertex = 0;
    fets.active_vertex_ratio = 0;
    fets.unactive_vertex = n;
    fets.unactive_vertex_ratio = 1;
    fets.growing_rate = fets.push_workload;
    fets.push_workload = 0;
    fets.pull_workload = m;
    fets.push_workload_ratio = 0;
    fets.pull_workload_ratio = 1;
    fets.cur_avg_deg_active = 0;
    fets.cur_avg_deg_unactive = fets.avg_deg;
  }

  void set_first_fets(active_set_t &as, feature_t &fets, const config_t &conf,
                      double n, double m) {
    if (fets.use_root < 0) {
      fets.active_vertex = n;
      fets.unactive_vertex = 0;
      fets.push_workload = m;
      fets.pull_workload = 0;
    } else {
      fets.active_vertex = fets.first_active;
      fets.unactive_vertex = n - fets.first_active;
      fets.push_workload = fets.first_workload;
      fets.pull_workload = m - fets.first_workload;
    }
    as.known_active = (long long)fets.active_vertex;
    as.known_inactive = (long long)fets.unactive_vertex;
    fets.active_vertex_ratio = DIV(fets.active_vertex, n);
    fets.unactive_vertex_ratio = DIV(fets.unactive_vertex, n);
    fets.push_workload_ratio = DIV(fets.push_workload, m);
    fets.pull_workload_ratio = DIV(fets.pull_workload, m);
    fets.cur_avg_deg_active = DIV(fets.push_workload, fets.active_vertex);
    fets.cur_avg_deg_unactive = DIV(fets.pull_workload, fets.unactive_vertex);
    fets.cur_max_deg_active = fets.push_workload;
    fets.cur_max_deg_unactive = fets.avg_deg;
  }

  void set_fets(active_set_t &as, stat_t &stat, feature_t &fets,
                const config_t &conf, double n, double m) {
    as.known_active = stat.active_num();
    as.known_inactive = stat.unactive_num();
    fets.active_vertex = (double)stat.active_num();
    fets.unactive_vertex = (double)stat.unactive_num();
    fets.active_vertex_ratio = DIV(fets.active_vertex, n);
    fets.unactive_vertex_ratio = DIV(fets.unactive_vertex, n);
    bool done = fets.active_vertex <= 0 ||
                (fets.pattern == Idem && fets.unactive_vertex <= 0);
    if (done) {
      if (conf.conf_qmode == Cached) {
        // keep the Cached cursor where a fused run would leave it
        as.queue.host_sizes[as.queue.cursor + 1] = 0;
        as.swap_queues();
      }
      as.halt_host();
    }
    fets.growing_rate = DIV((double)stat.push_workload(), fets.growing_rate);
    fets.push_workload = (double)stat.push_workload();
    fets.pull_workload = (double)stat.pull_workload();
    fets.push_workload_ratio = DIV(fets.push_workload, m);
    fets.pull_workload_ratio = DIV(fets.pull_workload, m);
    fets.cur_avg_deg_active = DIV(fets.push_workload, fets.active_vertex);
    fets.cur_avg_deg_unactive = DIV(fets.pull_workload, fets.unactive_vertex);
    fets.cur_max_deg_active = (double)stat.max_deg_active();
    fets.cur_max_deg_unactive = (double)stat.max_deg_unactive();
  }
};
