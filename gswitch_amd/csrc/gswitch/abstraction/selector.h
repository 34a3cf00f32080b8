// selector_t: the rule-based per-level variant choice.
//
// Decision logic follows selector_t (src/abstraction/selector.cuh:20-304):
// fusion triggers for idempotent algorithms (potential < 0.001 or three idle
// levels), fused levels run Push/Queue/WM, otherwise direction, frontier
// format and load balance come from the trained trees (now node tables,
// model/trees_table.h) on the same feature ratios, followed by the forced
// overrides. Class -> LB mapping is the reference's: 0 -> WM, 1 -> CM,
// >= 2 -> STRICT (selector.cuh:126-131). The frontier target status follows
// from (direction, from-all, to-all) as at selector.cuh:292-302.
#pragma once
#include "../model/mlp.h"
#include "../model/trees_table.h"
#include "config.h"
#include "features.h"
#include "statistics.h"

struct selector_t {
  void select(const stat_t &stats, feature_t &fets, config_t &conf) {
    select_based_on_model(stats, fets, conf);
  }

  static int tree(const gsw_model::tree_t &t, const double *fv) {
    return gsw_model::eval_tree(t, fv);
  }

  void trigger_fusion(const feature_t &fets, config_t &conf) {
    if (!conf.conf_fuse_inspect) {
      const double potential = fets.push_workload_ratio + fets.pull_workload_ratio;
      if (potential < 0.001) {
        conf.conf_fuse_inspect = true;
        conf.conf_switch_to_fusion = true;
      }
      if (fets.growing_rate > 2) conf.conf_idle = 0;
      if (conf.conf_idle >= 3) {
        conf.conf_fuse_inspect = true;
        conf.conf_switch_to_fusion = true;
        conf.conf_idle = 0;
      }
    }
  }

  void fusion(const feature_t &fets, config_t &conf) {
    conf.conf_dir = Push;
    conf.conf_asfmt = Queue;
    conf.conf_cm = Interleave;
    conf.conf_lb = WM;
    conf.conf_fuse_inspect = true;
    conf.conf_centric = fets.centric;
    conf.conf_fromall = fets.fromall;
    conf.conf_toall = fets.toall;
  }

  static void dir_features(const feature_t &fets, double *v) {
    const double *fv = fets.fv;
    v[0] = fv[8] / (fv[7] + 0.1);
    v[1] = fv[10] / (fv[9] + 0.1);
    v[2] = fv[16] / (fv[15] + 0.1);
    v[3] = fv[18] / (fv[17] + 0.1);
    v[4] = fv[9] / (fv[15] + 0.1);
    v[5] = fv[10] / (fv[16] + 0.1);
    v[6] = fv[9] / (fv[17] + 0.1);
    v[7] = fv[10] / (fv[18] + 0.1);
  }

  static LB lb_of_class(int l) { return l == 0 ? WM : (l == 1 ? CM : STRICT); }

  void select_idem(feature_t &fets, config_t &conf) {
    using namespace gsw_model;
    double v[16];
    dir_features(fets, v);
    conf.conf_dir = tree(k_select_dir_idem, v) == 1 ? Pull : Push;
    if (fets.push_workload_ratio * 5 < fets.pull_workload_ratio) conf.conf_dir = Push;
    if (conf.force.has_direction) conf.conf_dir = conf.force.d;

    if (conf.conf_dir == Pull) {
      conf.conf_asfmt = Bitmap;
      if (fets.fv[12] < 0.01) conf.conf_asfmt = Queue;
    } else {
      v[0] = fets.fv[11];
      conf.conf_asfmt = tree(k_select_fmt, v) ? Bitmap : Queue;
    }
    if (conf.force.has_asfmt) conf.conf_asfmt = conf.force.asf;

    if (conf.conf_dir == Pull) {
      conf.conf_lb = TM;
    } else {
      int k = 0;
      for (int i = 0; i < 7; ++i) v[k++] = fets.fv[i];
      for (int i = 7; i < 18; i += 2) v[k++] = fets.fv[i];
      conf.conf_lb = lb_of_class(tree(k_select_lb_idem, v));
    }
    if (conf.force.has_lb) conf.conf_lb = conf.force.lb;
  }

  void select_ASSO(feature_t &fets, config_t &conf) {
    using namespace gsw_model;
    double v[16];
    dir_features(fets, v);
    conf.conf_dir = tree(k_select_dir_asso, v) == 1 ? Pull : Push;
    if (conf.force.has_direction) conf.conf_dir = conf.force.d;
    if (conf.conf_fromall && conf.conf_toall) {
      // The reference asks k_select_dir_asso_backup here. On MI355X a
      // from-all, to-all push pays one device-scope atomic per edge (about
      // 25 random atomics per ns), where the pull reads every in-edge once
      // and writes each target plainly: on R-MAT 26 the tree picks the push
      // and PageRank takes 179 ms per iteration. Pull whatever the tree
      // says; the results do not depend on the direction.
      conf.conf_dir = Pull;
      if (conf.force.has_direction) conf.conf_dir = conf.force.d;
    }

    v[0] = (conf.conf_dir == Push) ? fets.fv[11] : fets.fv[12];
    conf.conf_asfmt = tree(k_select_fmt, v) ? Bitmap : Queue;
    if (fets.fromall && fets.toall && !conf.conf_fixed) conf.conf_asfmt = Bitmap;
    if (conf.force.has_asfmt) conf.conf_asfmt = conf.force.asf;

    int k = 0;
    for (int i = 0; i < 7; ++i) v[k++] = fets.fv[i];
    int l;
    const bool both = conf.conf_fromall && conf.conf_toall && !conf.conf_fixed;
    if (conf.conf_dir == Pull) {
      if (both) {
        l = tree(k_select_lb_asso_backward_backup, v);
      } else {
        for (int i = 8; i <= 18; i += 2) v[k++] = fets.fv[i];
        l = tree(k_select_lb_asso_backward, v);
      }
    } else {
      if (both) {
        l = tree(k_select_lb_asso_backward_backup, v);
      } else {
        for (int i = 7; i < 18; i += 2) v[k++] = fets.fv[i];
        l = tree(k_select_lb_asso_forward, v);
      }
    }
    conf.conf_lb = lb_of_class(l);
    if (conf.force.has_lb) conf.conf_lb = conf.force.lb;
  }

  void select_Mono(feature_t &fets, config_t &conf) {
    using namespace gsw_model;
    conf.conf_dir = Push;
    if (conf.force.has_direction) conf.conf_dir = conf.force.d;
    conf.conf_asfmt = Queue;
    if (conf.force.has_asfmt) conf.conf_asfmt = conf.force.asf;
    double v[16];
    int k = 0;
    for (int i = 0; i < 7; ++i) v[k++] = fets.fv[i];
    for (int i = 7; i < 18; i += 2) v[k++] = fets.fv[i];
    conf.conf_lb = lb_of_class(tree(k_select_lb_asso_forward, v));
    if (conf.force.has_lb) conf.conf_lb = conf.force.lb;
  }

  // Move TM/WM/CM to the edge-balanced STRICT when the frontier holds a
  // vertex whose degree exceeds conf.skew_limit. Idempotent pulls (early
  // exit per target) keep their choice.
  // From-all pulls over a bitmap (PageRank-like) go to TWOD instead: its
  // source slices are pulled one launch each, with rows binned by length.
  static void skew_guard(const feature_t &fets, config_t &conf) {
    if (conf.force.has_lb || conf.skew_limit <= 0) return;
    if (conf.conf_dir == Pull && conf.conf_pruning) return;
    const double maxd = (conf.conf_dir == Push) ? fets.cur_max_deg_active
                                                : fets.cur_max_deg_unactive;
    if (maxd <= conf.skew_limit) return;
    if (conf.conf_dir == Pull && conf.conf_fromall && conf.conf_asfmt == Bitmap)
      conf.conf_lb = TWOD;
    else if (conf.conf_lb == TM || conf.conf_lb == WM || conf.conf_lb == CM)
      conf.conf_lb = STRICT;
  }

  // The learned selector's class for this level. Entering fused mode
  // builds the queue from this level's bitmaps; leaving it happens after a
  // fused expand (the next level then runs the inspector), so a fused level
  // given a standalone class keeps the fused form with the class's LB.
  static void select_learned(config_t &conf) {
    conf.mlp_applied = conf.mlp_class;
    if (conf.mlp_pr_classes) {
      // one from-all pull for every iteration of the run
      const gsw_mlp::pr_variant_t p = gsw_mlp::pr_variant_of_class(conf.mlp_class);
      conf.conf_fuse_inspect = conf.conf_fusion = false;
      conf.conf_dir = Pull;
      conf.conf_asfmt = Bitmap;
      conf.conf_sorted = false;
      conf.conf_lb = p.lb;
      conf.mp_pull = p.pull_kernel != 1;
      conf.pb_pull = p.pull_kernel == 2;
      if (conf.force.has_direction) conf.conf_dir = conf.force.d;
      if (conf.force.has_asfmt) conf.conf_asfmt = conf.force.asf;
      if (conf.force.has_lb) conf.conf_lb = conf.force.lb;
      set_target(conf);
      return;
    }
    const gsw_mlp::variant_t v = gsw_mlp::variant_of_class(conf.mlp_class);
    if (v.fused || conf.conf_fuse_inspect) {
      if (!conf.conf_fuse_inspect) {
        conf.conf_switch_to_fusion = true;
        conf.mlp_entered_fused = true;
      }
      conf.conf_fuse_inspect = true;
      conf.conf_dir = Push;
      conf.conf_asfmt = Queue;
      conf.conf_sorted = false;
    } else {
      conf.conf_dir = v.dir;
      conf.conf_asfmt = v.fmt;
      conf.conf_sorted = v.sorted;
    }
    // the fusion mode flag follows the choice: with a stepping window, a
    // drained fused frontier must re-filter every vertex and leave fusion
    // before the window advances (active_set_t::finish)
    conf.conf_fusion = conf.conf_fuse_inspect;
    conf.conf_lb = v.lb;
    if (conf.force.has_direction) conf.conf_dir = conf.force.d;
    if (conf.force.has_asfmt) conf.conf_asfmt = conf.force.asf;
    if (conf.force.has_lb) conf.conf_lb = conf.force.lb;
    if (conf.force.has_sorted) conf.conf_sorted = conf.force.sorted;
    set_target(conf);
  }

  static void set_target(config_t &conf) {
    if (conf.conf_dir == Push)
      conf.conf_target = conf.conf_fromall ? All : Active;
    else
      conf.conf_target = conf.conf_toall ? All : Inactive;
  }

  void select_based_on_model(const stat_t &stats, feature_t &fets,
                             config_t &conf) {
    if (fets.centric == EC) {
      conf.conf_target = Active;
      conf.conf_asfmt = (fets.active_vertex_ratio > 0.8) ? Bitmap : Queue;
      if (conf.force.has_asfmt) conf.conf_asfmt = conf.force.asf;
      conf.conf_fuse_inspect = false;
      return;
    }
    // from-all & to-all: keep the first configuration for every level
    if (conf.conf_toall && conf.conf_fromall && !conf.conf_fixed) {
      conf.conf_inherit = true;
      return;
    }
    conf.conf_centric = fets.centric;
    conf.conf_fromall = fets.fromall;
    conf.conf_toall = fets.toall;

    if (conf.policy_mlp && conf.mlp_class >= 0) {
      select_learned(conf);
      return;
    }
    if (fets.pattern == Idem)
      trigger_fusion(fets, conf);
    else
      conf.conf_fuse_inspect = conf.conf_fusion;
    if (conf.force.has_fusion) {
      const bool want = conf.force.fusion == Fused;
      if (want && !conf.conf_fuse_inspect) conf.conf_switch_to_fusion = true;
      conf.conf_fuse_inspect = want;
    }
    if (conf.conf_fuse_inspect) {
      fusion(fets, conf);
      // forced variants may pin the LB of fused levels too
      if (conf.force.has_lb) conf.conf_lb = conf.force.lb;
      if (conf.force.has_sorted) conf.conf_sorted = conf.force.sorted;
      skew_guard(fets, conf);
      set_target(conf);
      return;
    }
    if (fets.pattern == Idem)
      select_idem(fets, conf);
    else if (fets.pattern == ASSO)
      select_ASSO(fets, conf);
    else
      select_Mono(fets, conf);
    if (conf.force.has_sorted) conf.conf_sorted = conf.force.sorted;
    skew_guard(fets, conf);
    set_target(conf);
  }
};
