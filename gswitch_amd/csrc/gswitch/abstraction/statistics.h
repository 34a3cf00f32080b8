// Inspector counters (stat_t, src/abstraction/statistics.cuh:9-55).
// Six device counters reduced by the inspector: active, inactive,
// push_workload (sum of out-degrees of active vertices), pull_workload (sum
// of in-degrees of inactive ones), max degree of an active / an inactive
// vertex. Counters are int64 (the reference's int32 overflows past 2^31
// scanned edges) and the maxima are true maxima (the reference block-sums
// them, inspector.cuh:101-109).
//
// Device layout: `dg_runtime_info[0..8)` holds the reduced counters the host
// copies back; after it come STAT_SLOTS partial slots, one 128-byte line
// each, that the counting passes add into (workgroup b uses slot
// b % STAT_SLOTS). Atomics from every workgroup onto one line serialise at
// about 12 ns each on MI355X (2048 workgroups x 6 counters = 150 us), while
// the same atomics spread over 64 lines cost nothing measurable. The
// one-workgroup finish kernel (k_stat_finish) sums the slots into the
// reduced counters and clears them, so the slots are zero between passes.
#pragma once
#include "../model/class_tree.h"
#include "../model/mlp.h"
#include "../utils/runtime.h"

constexpr int STAT_SLOTS = 64;   // partial-counter slots
constexpr int STAT_STRIDE = 16;  // long longs per slot (one 128-byte line)

struct stat_t {
  enum { N = 8, SLOT_BASE = 16, TOTAL = SLOT_BASE + STAT_SLOTS * STAT_STRIDE };
  long long *dg_runtime_info = nullptr;
  long long *h_runtime_info = nullptr;
  int size = 6;

  // history (wall-clock ms per phase)
  double last_time = 0, avg_time = 0;
  double last_filter_time = 0, last_expand_time = 0;
  double avg_filter_time = 0, avg_expand_time = 0;
  int level = 0;

  // learned selector (mlp.w == nullptr: rule trees only): the network the
  // counting passes run after their reduction, and a host copy of its
  // weights for levels decided without a counting pass
  gsw_mlp::sel_dev_t mlp;
  const float *mlp_host_w = nullptr;
  // class-tree selector (selector 2): evaluated on the host from the same
  // 12 features; mlp.w stays null, so no counting pass runs the network
  const gsw_tree::table_t *tree = nullptr;
  bool learned() const { return mlp.w != nullptr || tree != nullptr; }

  void build(int _size = 6) {
    size = _size;
    if (!dg_runtime_info) {
      H_ERR(hipMalloc((void **)&dg_runtime_info, sizeof(long long) * TOTAL));
      H_ERR(hipHostMalloc((void **)&h_runtime_info, sizeof(long long) * N));
    }
    for (int i = 0; i < N; ++i) h_runtime_info[i] = 0;
    clean();
  }
  void release() {
    if (dg_runtime_info) hipFree(dg_runtime_info);
    if (h_runtime_info) hipHostFree(h_runtime_info);
    dg_runtime_info = nullptr;
    h_runtime_info = nullptr;
  }
  // new run: history cleared, rule trees until a network is installed
  void reset() {
    last_time = avg_time = 0;
    last_filter_time = last_expand_time = 0;
    avg_filter_time = avg_expand_time = 0;
    level = 0;
    mlp = gsw_mlp::sel_dev_t{};
    mlp_host_w = nullptr;
    tree = nullptr;
  }
  // partial slots the counting passes add into
  long long *slots() const { return dg_runtime_info + SLOT_BASE; }
  // enqueue the device->host copy; the caller synchronises
  void collect_async(int num = N) {
    H_ERR(hipMemcpyAsync(h_runtime_info, dg_runtime_info,
                         sizeof(long long) * num, hipMemcpyDeviceToHost,
                         gsw::stream()));
  }
  void collect(int num = N) {
    collect_async(num);
    gsw::sync();
  }
  // zero the reduced counters and every slot
  void clean() { gsw::fill_async<long long>(dg_runtime_info, 0, TOTAL); }

  long long active_num() const { return h_runtime_info[0]; }
  long long unactive_num() const { return h_runtime_info[1]; }
  long long push_workload() const { return h_runtime_info[2]; }
  long long pull_workload() const { return h_runtime_info[3]; }
  long long max_deg_active() const { return h_runtime_info[4]; }
  long long max_deg_unactive() const { return h_runtime_info[5]; }
};
