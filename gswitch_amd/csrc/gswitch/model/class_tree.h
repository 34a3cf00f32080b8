// Variant-class decision trees over the learned selector's 12 features.
//
// The reference's rule selector is a set of decision trees trained offline
// from per-level records (src/model/*.h, README.md:13). A class tree is the
// same idea retrained for this engine on gfx950: one tree per application
// that maps the 12 selector features (gsw_mlp::features) straight to one of
// the 30 variant classes, fitted to the engine's own exhaustive per-level
// timings (gswitch_amd/learn.py fit-trees). Node i is internal when
// feat[i] >= 0 (go to left[i] if x[feat[i]] <= thr[i], else right[i]) and
// a leaf otherwise (class cls[i]). Evaluated on the host, where the
// counters of every inspected level are read anyway.
#pragma once
#include <cstdint>
#include <vector>

namespace gsw_tree {

struct table_t {
  std::vector<int32_t> feat, left, right, cls;
  std::vector<float> thr;

  bool empty() const { return feat.empty(); }
  // class of feature vector x (12 floats)
  int eval(const float *x) const {
    int i = 0;
    for (int steps = 0; feat[i] >= 0 && steps < (int)feat.size(); ++steps)
      i = (x[feat[i]] <= thr[i]) ? left[i] : right[i];
    return cls[i];
  }
};

}  // namespace gsw_tree
