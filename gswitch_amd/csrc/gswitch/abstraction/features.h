// Selector features (feature_t, src/abstraction/features.cuh:11-179).
// fv[0..20] has the reference's layout: graph (n, m, avg/std/range degree,
// Gini, entropy), workload (active/inactive counts, push/pull workloads and
// their ratios), current degrees (avg/max of active and inactive vertices)
// and history (last/avg filter and expand time). cross() folds the
// active and inactive sides together for from-all / to-all algorithms.
#pragma once
#include <fstream>
#include <sstream>
#include <string>

#include "../utils/common.h"

// Destination of feature_t::record(): a file the records are appended to
// ("" = off) and how many were written since it was set.
struct record_sink_t {
  std::string path;
  int written = 0;
};
inline record_sink_t &record_sink() {
  static record_sink_t s;
  return s;
}

struct feature_t {
  // Algorithm
  Centric centric = VC;
  Computation pattern = Idem;
  int use_root = -1;
  bool fromall = false;
  bool toall = false;

  // Dataset
  double nvertexs = 0;
  double nedges = 0;
  double avg_deg = 0, std_deg = 0, max_deg = 0, range = 0, GI = 0, Her = 0;
  double avg_edata = 0;

  // Runtime
  double active_vertex = 0, unactive_vertex = 0;
  double active_vertex_ratio = 0, unactive_vertex_ratio = 0;
  double push_workload = 0, pull_workload = 0;
  double push_workload_ratio = 0, pull_workload_ratio = 0;
  double cur_max_deg_active = 0, cur_avg_deg_active = 0;
  double cur_max_deg_unactive = 0, cur_avg_deg_unactive = 0;

  // History
  int level = 0;
  double last_time = 0, avg_time = 0;
  double last_filter_time = 0, last_expand_time = 0;
  double avg_filter_time = 0, avg_expand_time = 0;
  double growing_rate = 0;
  double first_workload = 0;
  double first_active = 1;  // roots of the first frontier (multi-root runs)

  // Architecture
  double cap = 0;
  int thread_lim = 1024;
  int register_lim = 0;
  int sm_num = 0;

  double fv[21] = {0};

  void flatten() {
    fv[0] = nvertexs;
    fv[1] = nedges;
    fv[2] = avg_deg;
    fv[3] = std_deg;
    fv[4] = range;
    fv[5] = GI;
    fv[6] = Her;
    fv[7] = active_vertex;
    fv[8] = unactive_vertex;
    fv[9] = push_workload;
    fv[10] = pull_workload;
    fv[11] = active_vertex_ratio;
    fv[12] = unactive_vertex_ratio;
    fv[13] = push_workload_ratio;
    fv[14] = pull_workload_ratio;
    fv[15] = cur_avg_deg_active;
    fv[16] = cur_avg_deg_unactive;
    fv[17] = cur_max_deg_active;
    fv[18] = cur_max_deg_unactive;
    fv[19] = DIV(last_filter_time, avg_filter_time);
    fv[20] = DIV(last_expand_time, avg_expand_time);
    cross();
  }

  // One training record in the reference's format (feature_t::record,
  // src/abstraction/features.cuh:86-94 with json_t::dump,
  // src/utils/json.cuh:31-44): keys "0".."20" = fv, "21" = last filter
  // time, "22" = last expand time (ms), values printed as an ostream prints
  // a double, the object appended to the file; at most 100 records per
  // file setting, as json_t stops after 100 dumps.
  void record() {
    record_sink_t &k = record_sink();
    if (k.path.empty() || k.written >= 100) return;
    flatten();
    std::ostringstream o;
    o << "{\n";
    for (int i = 0; i < 23; ++i) {
      const double x = i < 21 ? fv[i] : (i == 21 ? last_filter_time : last_expand_time);
      if (i > 0) o << ",\n";
      o << "\"" << i << "\": " << x;
    }
    o << "\n}\n";
    std::ofstream out(k.path, std::ios::app);
    out << o.str();
    ++k.written;
  }

  void cross() {
    const double tot = active_vertex + unactive_vertex;
    const double mixed =
        tot == 0 ? 0
                 : (cur_avg_deg_active * active_vertex +
                    cur_avg_deg_unactive * unactive_vertex) / tot;
    if (fromall) {
      fv[7] += unactive_vertex;
      fv[9] += pull_workload;
      fv[11] += unactive_vertex_ratio;
      fv[13] += pull_workload_ratio;
      fv[15] += mixed;
      fv[17] = MAX(cur_max_deg_unactive, cur_max_deg_active);
    }
    if (toall) {
      fv[8] += active_vertex;
      fv[10] += push_workload;
      fv[12] += active_vertex_ratio;
      fv[14] += push_workload_ratio;
      fv[16] += mixed;
      fv[18] = MAX(cur_max_deg_active, cur_max_deg_unactive);
    }
    active_vertex = fv[7];
    unactive_vertex = fv[8];
    push_workload = fv[9];
    pull_workload = fv[10];
    active_vertex_ratio = fv[11];
    unactive_vertex_ratio = fv[12];
    push_workload_ratio = fv[13];
    pull_workload_ratio = fv[14];
    cur_avg_deg_active = fv[15];
    cur_avg_deg_unactive = fv[16];
    cur_max_deg_active = fv[17];
    cur_max_deg_unactive = fv[18];
  }

  void architecture_features(int device = 0) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
      cap = prop.major + prop.minor / 10.0;
      register_lim = prop.regsPerBlock / 512;
      thread_lim = prop.maxThreadsPerBlock;
      sm_num = prop.multiProcessorCount;
    }
  }

  // Graph-level features from the degree statistics (Fake<VC>::func,
  // src/data_structures/graph.cuh:185-200).
  template <typename Attr>
  void set_graph(double n, double m_undirected, const Attr &a,
                 double mean_weight = 0) {
    nvertexs = n;
    nedges = m_undirected;
    avg_deg = a.avg_deg;
    std_deg = a.std_deg;
    max_deg = a.max_deg;
    range = a.range;
    GI = a.GI;
    Her = a.Her;
    avg_edata = mean_weight;
  }
};
