// GSWITCH-compatible traversal engine for MI355X (gfx950, wave64).
//
// Umbrella header with the reference's public glue (src/gswitch.h:19-108):
// init_conf / reset_conf, plus active_set_t::finish and p_jump. Application
// code keeps the reference's shape: declare the actors, subclass Functor,
// build the graph and the active set, and run the level loop
//   inspect -> finish -> select -> filter -> update_level -> expand -> finish.
#pragma once
#include "abstraction/config.h"
#include "abstraction/executor.h"
#include "abstraction/features.h"
#include "abstraction/inspector.h"
#include "abstraction/selector.h"
#include "abstraction/statistics.h"
#include "data_structures/active_set.h"
#include "data_structures/functor.h"
#include "data_structures/graph.h"
#include "kernels/misc.h"
#include "model/trees_table.h"
#include "utils/cmdline.h"
#include "utils/fileio.h"

// init_conf (src/gswitch.h:19-63): initial fused/standalone choice from the
// select_fusion tree on (avg, std, range, GI, Her) of the degrees, disabled
// for from-all & to-all algorithms, forced override, SSSP window size
// 8 * mean_weight * 32 / avg_degree.
// A command line parsed by parse_cmd adds its pinned choices (--configs),
// its verbosity and its training-record file (--json).
template <typename Gr, typename F>
void init_conf(stat_t &stats, feature_t &fets, config_t &conf, Gr &g, F &f) {
  stats.build();
  fets.architecture_features();
  conf.conf_inherit = false;
  if (cmd_forces_anything(cmd_opt.ins) && !cmd_forces_anything(conf.force))
    conf.force = cmd_opt.ins;
  if (cmd_opt.verbose) conf.verbose = true;
  if (!cmd_opt.json.empty() && record_sink().path != cmd_opt.json) {
    record_sink().path = cmd_opt.json;
    record_sink().written = 0;
  }
  if (fets.use_root >= 0 && fets.first_active <= 1)
    fets.first_workload = g.root_degree(fets.use_root);
  if (fets.centric == VC) {
    fets.flatten();
    double fv[5];
    for (int i = 2; i < 7; ++i) fv[i - 2] = fets.fv[i];
    conf.conf_fuse_inspect =
        gsw_model::eval_tree(gsw_model::k_select_fusion, fv) != 0;
    if (fets.toall && fets.fromall) conf.conf_fuse_inspect = false;
    if (conf.force.has_fusion) conf.conf_fuse_inspect = conf.force.fusion == Fused;
    if (conf.conf_fuse_inspect) conf.conf_fusion = true;
    if (!conf.conf_fuse_inspect) conf.conf_qmode = conf.conf_qmode;
  }
  if (conf.conf_window) {
    f.data.window.enable = true;
    f.data.window.set_init_winsize(DIV(8 * g.mean_weight * 32, g.attr.avg_deg));
  }
}

template <typename Gr, typename F>
void reset_conf(stat_t &stats, feature_t &fets, config_t &conf, Gr &g, F &f) {
  conf.reset();
  conf.conf_fusion = false;
  init_conf(stats, fets, conf, g, f);
}

// Convergence test after a level (active_set_t::finish + notification_t::
// exit, src/data_structures/active_set.cuh:37-43, notification.cuh:73-89):
// with the SSSP window, a drained frontier only ends the run when no vertex
// still has a pending update; otherwise the window advances and the loop
// continues (in standalone mode if it was fused). `window_moved` tells a
// level loop that called this after an inspection to inspect again (with
// the new window) instead of expanding the drained frontier.
template <typename G, typename F>
bool active_set_t::finish(G &g, F &f, config_t &conf) {
  if (!converged) return false;
  if constexpr (std::is_same<G, device_graph_t<CSR, typename G::edge_t>>::value) {
    if (!conf.conf_window) return true;
    if (conf.conf_fusion) {
      hipLaunchKernelGGL((gsw::k_filter_all<G, F>),
                         dim3(gsw::grid_for(g.nvertexs)), dim3(gsw::BLOCK), 0,
                         gsw::stream(), g, f);
    }
    gsw::set_one_async<int32_t>(dg_ctrl + CTRL_FLAG, 0);
    hipLaunchKernelGGL((gsw::k_exit_any<G, F>), dim3(gsw::grid_for(g.nvertexs)),
                       dim3(gsw::BLOCK), 0, gsw::stream(), g, f,
                       dg_ctrl + CTRL_FLAG);
    H_ERR(hipGetLastError());
    H_ERR(hipMemcpyAsync(h_ctrl + CTRL_FLAG, dg_ctrl + CTRL_FLAG,
                         sizeof(int32_t), hipMemcpyDeviceToHost, gsw::stream()));
    gsw::sync();
    if (h_ctrl[CTRL_FLAG]) {
      converged = false;
      window_moved = true;
      f.data.window.reset();
      if (conf.conf_fusion) {
        conf.conf_fuse_inspect = false;
        conf.conf_switch_to_fusion = true;
      }
      return false;
    }
    return true;
  } else {
    return true;
  }
}

// Pointer jumping to a fixpoint (p_jump, src/utils/intrinsics.cuh:351-362):
// the functor by value as in the reference (it only carries device
// pointers), dg_flag one device int, h_flag one pinned host int.
template <typename F>
void p_jump(F f, int64_t nvertexs, int32_t *dg_flag, int32_t *h_flag) {
  for (;;) {
    gsw::set_one_async<int32_t>(dg_flag, 0);
    hipLaunchKernelGGL(gsw::k_pointer_jumping, dim3(gsw::grid_for(nvertexs)),
                       dim3(gsw::BLOCK), 0, gsw::stream(), (int *)f.data.dg_wa,
                       nvertexs, (int *)dg_flag);
    H_ERR(hipGetLastError());
    H_ERR(hipMemcpyAsync(h_flag, dg_flag, sizeof(int32_t),
                         hipMemcpyDeviceToHost, gsw::stream()));
    gsw::sync();
    if (*h_flag == 0) break;
  }
}
