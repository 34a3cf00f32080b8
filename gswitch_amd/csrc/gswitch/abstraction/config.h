// Per-level execution configuration chosen by the selector.
//
// Field names follow config_t (src/abstraction/config.cuh:9-123). Additions:
// conf_sorted (sorted vs unsorted queue frontier), the TWC bin thresholds,
// and `force`, the per-run form of the reference's `--configs=Dir-Fmt-LB-
// Fusion` override (src/utils/cmdline.cuh:51-94) that the selector applies
// after its own decision.
#pragma once
#include <iostream>

#include "../utils/common.h"

struct ins_t {
  bool has_direction = false, has_asfmt = false, has_lb = false;
  bool has_fusion = false, has_sorted = false;
  Direction d = Push;
  ASFmt asf = Queue;
  LB lb = WM;
  Fusion fusion = Standalone;
  bool sorted = false;
};

struct config_t {
  Direction conf_dir = Push;
  Status conf_target = Active;
  CullMode conf_cm = Interleave;
  ASFmt conf_asfmt = Queue;
  LB conf_lb = WM;
  Centric conf_centric = VC;
  QueueMode conf_qmode = Normal;

  bool conf_fromall = false;
  bool conf_toall = false;
  bool conf_fixed = false;
  bool conf_inherit = false;
  bool conf_2d_partition = false;

  bool conf_first_round = true;
  bool conf_switch_to_fusion = false;
  bool conf_switch_to_standalone = false;
  bool conf_pruning = false;
  bool conf_once = false;
  bool conf_ignore_u_state = false;
  bool conf_fuse_inspect = false;
  bool conf_lite = false;
  bool conf_window = false;
  int wm_vpw = 64;  // frontier items per wavefront of the WM push
  // TM pulls take each target's first in-edge from the per-vertex copy
  bool pull_first = false;
  bool conf_compensation = false;
  bool conf_super_fusion = false;
  bool conf_fusion = false;
  bool conf_sorted = false;
  int conf_idle = 0;
  int ctanum = 256;
  int thdnum = gsw::BLOCK;
  // TWC bins: degree < twc_small -> thread, < twc_large -> wavefront,
  // otherwise workgroup.
  int twc_small = 32;
  int twc_large = 2048;
  // skew guard: a TM/WM/CM choice moves to STRICT when a frontier vertex has
  // more edges than this (one lane / wavefront / workgroup would otherwise
  // walk a hub alone); 0 disables it. Forced LBs are never changed.
  int skew_limit = 4096;
  // from-all bitmap pulls use the merge-path kernel (k_pull_mp*) instead of
  // the segmented-reduction STRICT/TWOD kernels
  bool mp_pull = true;
  // from-all pulls of unweighted graphs through the propagation-blocked
  // layout (kernels/pb.h) instead of the TWOD slices
  bool pb_pull = false;
  // learned selector: every level runs the variant class the MLP picks
  // (mlp_class, written by the inspector / the fused-level selection)
  bool policy_mlp = false;
  int mlp_class = -1;
  // the class is one of the from-all pull kernels (gsw_mlp::pr_variant_of_class),
  // chosen once, on the first level
  bool mlp_pr_classes = false;
  int mlp_applied = -1;  // class of the latest learned selection applied
  // fused levels under the learned selector: the frontier size when the
  // network last chose a fused class (re-inspection waits for growth)
  bool mlp_entered_fused = false;
  long long mlp_fused_ref = 0;
  // set by an expand that already inspected the next level (TM pull of a
  // level-synchronous functor): the inspector then only reads the counters
  bool conf_pre_inspected = false;
  // every level has every vertex active (a from-all, to-all run whose filter
  // never deactivates one, e.g. PageRank with eps = 0): the inspector runs
  // the filter pass but reads no counters back, so levels need no host sync
  bool counts_known = false;
  // with counts_known, a propagation-blocked expand applies the functor's
  // filter to every target after combining its messages (what the next
  // level's inspection would do first; such a filter reads no level);
  // filter_done tells that inspection
  bool pb_fuse_filter = false;
  // the functor's filter leaves every WA at zero, its combine's identity
  // (PageRank): a gather that applies the filter starts its LDS
  // accumulators of whole bins from zero instead of reading the WA back
  bool wa_zero = false;
  bool filter_done = false;
  ins_t force;
  // one line per expand on stdout (the reference's --verbose)
  bool verbose = false;

  void reset() {
    conf_first_round = true;
    conf_switch_to_fusion = false;
    conf_switch_to_standalone = false;
    conf_fuse_inspect = false;
    conf_pre_inspected = false;
    conf_idle = 0;
  }

  __host__ __device__ __forceinline__ Status want() const { return conf_target; }
  __host__ __device__ __forceinline__ bool pruning() const { return conf_pruning; }
  __host__ __device__ __forceinline__ bool fuse_inspect() const {
    return conf_fuse_inspect;
  }
  __host__ __device__ __forceinline__ bool ignore_u_state() const {
    return conf_ignore_u_state;
  }

  static const char *lb_name(LB lb) {
    switch (lb) {
      case TM: return "TM";
      case WM: return "WM";
      case CM: return "CM";
      case STRICT: return "STRICT";
      case TWOD: return "2D-partition";
      case ELB: return "ELB";
      case TWC: return "TWC";
    }
    return "Unknown";
  }

  void show_hints(std::ostream &out = std::cout) const {
    out << "[" << (conf_dir == Push ? "Push|" : "Pull|")
        << (conf_asfmt == Queue ? (conf_sorted ? "SortedQueue|" : "Queue|")
                                : "Bitmap|")
        << lb_name(conf_lb) << (conf_fuse_inspect ? "|Fused" : "|Standalone")
        << "] ";
  }
};

// Parse "Dir-Fmt-LB-Fusion" (any field may be "x"), as prase_configs does
// (src/utils/cmdline.cuh:51-94); Fmt additionally accepts "SortedQueue".
inline ins_t parse_configs(const std::string &s) {
  ins_t ins;
  std::string parts[4];
  int k = 0;
  std::string cur;
  for (char c : s) {
    if (c == '-') {
      if (k < 4) parts[k++] = cur;
      cur.clear();
    } else {
      cur += c;
    }
  }
  if (k < 4) parts[k++] = cur;
  if (parts[0] == "Push") { ins.has_direction = true; ins.d = Push; }
  else if (parts[0] == "Pull") { ins.has_direction = true; ins.d = Pull; }
  if (parts[1] == "Queue") { ins.has_asfmt = true; ins.asf = Queue; ins.has_sorted = true; ins.sorted = false; }
  else if (parts[1] == "SortedQueue") { ins.has_asfmt = true; ins.asf = Queue; ins.has_sorted = true; ins.sorted = true; }
  else if (parts[1] == "Bitmap") { ins.has_asfmt = true; ins.asf = Bitmap; }
  const std::string &l = parts[2];
  if (l == "TM") { ins.has_lb = true; ins.lb = TM; }
  else if (l == "WM") { ins.has_lb = true; ins.lb = WM; }
  else if (l == "CM") { ins.has_lb = true; ins.lb = CM; }
  else if (l == "STRICT") { ins.has_lb = true; ins.lb = STRICT; }
  else if (l == "TWC") { ins.has_lb = true; ins.lb = TWC; }
  else if (l == "TWOD") { ins.has_lb = true; ins.lb = TWOD; }
  if (parts[3] == "Fused") { ins.has_fusion = true; ins.fusion = Fused; }
  else if (parts[3] == "Standalone") { ins.has_fusion = true; ins.fusion = Standalone; }
  return ins;
}
