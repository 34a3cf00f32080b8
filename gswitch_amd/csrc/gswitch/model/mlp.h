// Learned variant selector: a 12 -> 64 -> 64 -> 30 FP32 MLP.
//
// No counterpart in the reference, which selects with decision trees
// trained offline from feature dumps (src/abstraction/features.cuh:86-104,
// src/model/select_*.h). Here the network regresses, for each of the 30
// expand variants, the log2 of its per-level cost relative to the fastest
// variant; the engine runs the valid variant with the smallest prediction.
//
// Parameter layout (6942 floats, row-major, input index first):
//   W1[12][64] b1[64] W2[64][64] b2[64] W3[64][30] b3[30]
//
// Variant classes: class = group * 5 + lb, lb in {TM, WM, CM, STRICT, TWC},
//   group 0 Push-Queue-Standalone      3 Push-Queue-Fused
//         1 Push-SortedQueue-Standalone 4 Pull-Queue-Standalone
//         2 Push-Bitmap-Standalone      5 Pull-Bitmap-Standalone
//
// Features (all FP32, identical on host and device):
//   0 log2(1+n)/24        4 Gini of the degrees    8 log2(1+push workload)/32
//   1 log2(1+m)/32        5 log2(1+level)/8        9 log2(1+pull workload)/32
//   2 log2(1+avg deg)/8   6 active / n            10 log2(1+push wl/active)/16
//   3 log2(1+std/avg)/4   7 inactive / n          11 log2(1+max active deg)/24
// where m counts directed CSR entries, "active" the frontier, "inactive"
// the vertices a pull would visit, and the workloads are the edge counts a
// push (out-edges of the frontier) or a pull (in-edges of the inactive
// vertices) would scan. The workloads enter as absolute log sizes so that
// the cost of a variant on a larger graph than any seen in training is an
// extrapolation along a known axis.
#pragma once
#include <cmath>
#include <cstdint>

#include "../utils/common.h"

namespace gsw_mlp {

constexpr int NIN = 12, NH = 64, NOUT = 30;
constexpr int OW1 = 0, OB1 = OW1 + NIN * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH,
              OW3 = OB2 + NH, OB3 = OW3 + NH * NOUT, NPARAM = OB3 + NOUT;
static_assert(NPARAM == 6942, "parameter count");
constexpr int NGROUP = 6, NLB = 5;
constexpr uint32_t ALL_CLASSES = (1u << NOUT) - 1;

struct variant_t {
  Direction dir;
  ASFmt fmt;
  bool sorted;
  bool fused;
  LB lb;
};

__host__ __device__ inline LB lb_of_index(int i) {
  return i == 0 ? TM : i == 1 ? WM : i == 2 ? CM : i == 3 ? STRICT : TWC;
}
__host__ __device__ inline int index_of_lb(LB lb) {
  switch (lb) {
    case TM: return 0;
    case WM: return 1;
    case CM: return 2;
    case TWC: return 4;
    default: return 3;  // STRICT and its edge-balanced relatives
  }
}

__host__ __device__ inline variant_t variant_of_class(int c) {
  const int g = c / NLB;
  variant_t v;
  v.lb = lb_of_index(c % NLB);
  v.dir = (g >= 4) ? Pull : Push;
  v.fmt = (g == 2 || g == 5) ? Bitmap : Queue;
  v.sorted = (g == 1);
  v.fused = (g == 3);
  return v;
}

inline int class_of(Direction dir, ASFmt fmt, bool sorted, bool fused, LB lb) {
  int g;
  if (dir == Push)
    g = fused ? 3 : (fmt == Bitmap ? 2 : (sorted ? 1 : 0));
  else
    g = (fmt == Bitmap) ? 5 : 4;
  return g * NLB + index_of_lb(lb);
}

// PageRank-like runs (from-all and to-all: every iteration pulls every
// vertex in the same way) choose once per run among the from-all pulls of a
// bitmap frontier; their classes are the network's first NPR outputs:
//   0 TM   1 WM   2 CM   3 STRICT, merge-path walk   4 STRICT, segmented
//   5 TWOD, length-binned source slices   6 TWOD, segmented slices
//   7 TWOD, propagation blocking
// (pull_kernel of each: the from-all pull kernel option of gsw_run_opts).
constexpr int NPR = 8;
constexpr uint32_t PR_CLASSES = (1u << NPR) - 1;
struct pr_variant_t {
  LB lb;
  int pull_kernel;
};
__host__ __device__ inline pr_variant_t pr_variant_of_class(int c) {
  switch (c) {
    case 0: return {TM, 2};
    case 1: return {WM, 2};
    case 2: return {CM, 2};
    case 3: return {STRICT, 0};
    case 4: return {STRICT, 1};
    case 5: return {TWOD, 0};
    case 6: return {TWOD, 1};
    default: return {TWOD, 2};
  }
}

// Connected components choose once per run, from the same first-level
// inputs (every vertex active), among the network's first NCC outputs:
//   0 sampled hooking over the CSR   1 edge-centric level loop, frontier
//   formats by the rule   2 edge-centric, bitmap   3 edge-centric, edge queue
constexpr int NCC = 4;
constexpr uint32_t CC_CLASSES = (1u << NCC) - 1;

// classes that make sense for an algorithm: a from-all and to-all one
// (PageRank-like) has no fused form, as init_conf keeps it standalone
inline uint32_t valid_mask(bool fromall, bool toall) {
  uint32_t m = ALL_CLASSES;
  if (fromall && toall) m &= ~(((1u << NLB) - 1) << (3 * NLB));
  return m;
}

// per-level counters the features come from
struct counts_t {
  float n, m, avg_deg, std_deg, gini;
  float level, active, inactive, push_wl, pull_wl, max_deg_active;
};

__host__ __device__ inline float safe_div(float a, float b) {
  return b > 0.f ? a / b : 0.f;
}

__host__ __device__ inline void features(const counts_t &c, float *x) {
  x[0] = log2f(1.f + c.n) / 24.f;
  x[1] = log2f(1.f + c.m) / 32.f;
  x[2] = log2f(1.f + c.avg_deg) / 8.f;
  x[3] = log2f(1.f + safe_div(c.std_deg, c.avg_deg)) / 4.f;
  x[4] = c.gini;
  x[5] = log2f(1.f + fmaxf(c.level, 0.f)) / 8.f;
  x[6] = safe_div(c.active, c.n);
  x[7] = safe_div(c.inactive, c.n);
  x[8] = log2f(1.f + c.push_wl) / 32.f;
  x[9] = log2f(1.f + c.pull_wl) / 32.f;
  x[10] = log2f(1.f + safe_div(c.push_wl, fmaxf(c.active, 1.f))) / 16.f;
  x[11] = log2f(1.f + c.max_deg_active) / 24.f;
}

// Host forward pass (first levels, whose features are known without an
// inspector pass); the same arithmetic as the device path, in FP32.
inline void forward_host(const float *w, const float *x, float *y) {
  float h1[NH], h2[NH];
  for (int j = 0; j < NH; ++j) {
    float a = w[OB1 + j];
    for (int k = 0; k < NIN; ++k) a += x[k] * w[OW1 + k * NH + j];
    h1[j] = a > 0.f ? a : 0.f;
  }
  for (int j = 0; j < NH; ++j) {
    float a = w[OB2 + j];
    for (int k = 0; k < NH; ++k) a += h1[k] * w[OW2 + k * NH + j];
    h2[j] = a > 0.f ? a : 0.f;
  }
  for (int j = 0; j < NOUT; ++j) {
    float a = w[OB3 + j];
    for (int k = 0; k < NH; ++k) a += h2[k] * w[OW3 + k * NOUT + j];
    y[j] = a;
  }
}

// argmin over the valid classes, ties to the lower class id; -1 if none
inline int argmin_valid(const float *y, uint32_t mask) {
  int best = -1;
  for (int j = 0; j < NOUT; ++j)
    if ((mask >> j) & 1u)
      if (best < 0 || y[j] < y[best]) best = j;
  return best;
}

// Device-side selector state handed to the selection kernel (by value).
// `w` = 0 disables inference.
struct sel_dev_t {
  const float *w = nullptr;
  float n = 0, m = 0, avg_deg = 0, std_deg = 0, gini = 0;
  float level = 0;
  uint32_t mask = ALL_CLASSES;
};

// On a fused level the next frontier's length is known without a pass. A
// frontier of at least n / INSPECT_FRACTION vertices that has also grown
// INSPECT_GROWTH-fold since the network last chose a fused class is handed
// to the next level's inspector (and its inference); otherwise the level
// stays fused with the current class.
constexpr double INSPECT_FRACTION = 64.0;
constexpr double INSPECT_GROWTH = 2.0;

}  // namespace gsw_mlp
