// SSSP stepping window (delta-stepping-like bucket of tentative distances).
//
// Follows window_t (src/data_structures/window.cuh:44-123): the window is
// [th, th + wsz); after each level the rule tree select_stepping on
// (vertices processed, estimated edges) shrinks it (x0.3, twice when the
// estimate is huge), inflates it (/0.1, capped at the initial size) or keeps
// it; when the window drains, reset() advances th by wsz. Only the amount of
// work per level depends on the window, never the final distances.
#pragma once
#include <vector>

#include "../model/trees_table.h"
#include "../utils/common.h"

struct window_t {
  double th = 0, wsz_lim = 0, wsz = 0;
  double shrink_rate = 0.3;
  double inflat_rate = 0.1;
  double reset_rate = 1;
  bool enable = false;

  bool too_large(double size) const { return size > 5 * 33679; }

  // `active` = vertices processed last level, `avg_deg` = their mean degree.
  void adjust(double active, double avg_deg, bool fusion) {
    if (fusion) {
      wsz = wsz_lim;
      return;
    }
    double fv[2] = {active, active * avg_deg};
    int l = gsw_model::eval_tree(gsw_model::k_select_stepping, fv);
    if (!enable) return;
    if (l == 1) {
      shrink();
      if (too_large(fv[1])) shrink();
    } else if (l == 0) {
      inflat();
    }
  }

  void set_init_winsize(double lim) {
    th = 0;
    wsz_lim = lim;
    wsz = lim;
  }
  void shrink() { wsz = wsz * shrink_rate; }
  void inflat() {
    wsz = wsz / inflat_rate;
    if (wsz > wsz_lim) wsz = wsz_lim;
  }
  void reset() {
    th = th + wsz;
    wsz = wsz_lim * reset_rate;
  }

  // Distances are compared as double: exact for every uint32 value.
  __host__ __device__ __forceinline__ bool in(double x) const {
    return !enable || (x < th + wsz && x >= th);
  }
  __host__ __device__ __forceinline__ bool less(double x) const { return x < th; }
  __host__ __device__ __forceinline__ bool greater(double x) const {
    return x >= th + wsz;
  }
};
