// executor_t: frontier generation and expand dispatch.
//
// filter() follows executor_t::filter + __filter (src/abstraction/
// executor.cuh:21-72, src/kernel_libs/filter.cuh:197-252): set the frontier
// format, compact bitmap -> queue when a queue is needed and it was not
// produced by a fused expand, swap the queues, size the grid and run the
// fusion compensation pass. expand() follows executor_t::expand
// (executor.cuh:74-148): one launch of the chosen {direction, LB, fusion}
// variant, timing history, and the fused -> standalone fallback when an
// expand takes more than 5x the average.
//
// Host<->device traffic per level: standalone levels sync once (the
// inspector's counters); fused levels sync once after the expand (the next
// queue's length, which also tells convergence). Queue lengths are otherwise
// known on the host from the inspector's counts.
#pragma once
#include <vector>

#include "../data_structures/active_set.h"
#include "../kernels/expand_pull.h"
#include "../kernels/expand_push.h"
#include "../kernels/filter.h"
#include "../kernels/misc.h"
#include "../kernels/pb.h"
#include "../kernels/twod.h"
#include "config.h"
#include "features.h"
#include "inspector.h"
#include "statistics.h"

// One record per expand, for logs, training data and roofline accounting.
struct level_record_t {
  int level;
  int dir, fmt, lb, fused, sorted;
  int kernel;              // pulls: 0 length-binned / merge-path, 1 segmented,
                           // 2 propagation blocking (pull_kernel in effect)
  long long frontier;      // input frontier length (queue) or n (bitmap)
  long long active, inactive;
  long long edges;         // edges inspected by the expand (device counter)
  long long verts;         // frontier vertices expanded
  float ms;                // expand time (hipEvent)
};

struct executor_t {
  bool profile = true;
  std::vector<level_record_t> log;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, int>> ev_pairs;  // per log entry
  int ev_next = 0;

  void clear_log() {
    log.clear();
    ev_pairs.clear();
    ev_next = 0;
  }

  void release() {
    for (auto e : ev_pool) hipEventDestroy(e);
    ev_pool.clear();
    clear_log();
  }

  hipEvent_t next_event() {
    if (ev_next == (int)ev_pool.size()) {
      hipEvent_t e;
      H_ERR(hipEventCreate(&e));
      ev_pool.push_back(e);
    }
    return ev_pool[ev_next++];
  }

  // Resolve event timings of all logged expands (syncs the stream).
  void finalize_log(active_set_t &as) {
    gsw::sync();
    for (size_t i = 0; i < log.size(); ++i) {
      if (ev_pairs[i].first < 0) continue;
      float ms = 0;
      H_ERR(hipEventElapsedTime(&ms, ev_pool[ev_pairs[i].first],
                                ev_pool[ev_pairs[i].second]));
      log[i].ms = ms;
    }
    if (!log.empty() && as.dg_work_base) {
      const int64_t rows = std::min<int64_t>((int64_t)log.size(), WORK_ROWS);
      hipLaunchKernelGGL(k_work_fold, dim3((rows * WORK_N + 255) / 256), dim3(256),
                         0, gsw::stream(), as.dg_work_base, rows);
      H_ERR(hipGetLastError());
      gsw::sync();
      std::vector<unsigned long long> w(rows * WORK_N);
      H_ERR(hipMemcpy(w.data(), as.dg_work_base,
                      sizeof(unsigned long long) * w.size(),
                      hipMemcpyDeviceToHost));
      for (size_t i = 0; i < log.size(); ++i) {
        const size_t r = std::min<size_t>(i, (size_t)rows - 1);
        log[i].edges = (long long)w[r * WORK_N + WORK_EDGES];
        log[i].verts = (long long)w[r * WORK_N + WORK_VERTS];
      }
    }
  }

  // ------------------------------------------------------------- filter
  template <typename G, typename F>
  void filter(active_set_t &as, G &g, F &f, stat_t &stat, feature_t &fets,
              config_t &conf) {
    const double s = stat.last_filter_time;
    as.set_fmt(conf.conf_asfmt);
    if (conf.conf_asfmt == Queue) {
      bool tofilter = !conf.conf_fuse_inspect;
      if (conf.conf_switch_to_fusion) {
        tofilter = true;
        conf.conf_switch_to_fusion = false;
      }
      if (conf.conf_first_round) tofilter = true;
      if (tofilter) {
        if (!(as.queue.mode == Cached && as.queue.traceback))
          run_filter(as, conf);
      }
      // a fused level of a non-idempotent algorithm de-duplicates its
      // pushes within the level only: start from a clean `visited`, also on
      // the level that switches into fusion (an earlier fused phase may have
      // left marks behind)
      if (conf.conf_fuse_inspect && !conf.conf_pruning)
        gsw::fill_async<uint64_t>(as.bitmap.visited.dg_bits, 0,
                                  as.bitmap.visited.n_words + 1);
      as.swap_queues();
    }
    conf.thdnum = gsw::BLOCK;
    if (!conf.conf_fuse_inspect) {
      const double wl = (conf.conf_dir == Push) ? fets.push_workload
                                                : fets.pull_workload;
      conf.ctanum = (int)MIN(256.0, wl / conf.thdnum / 2 + 1);
      if (conf.ctanum < 4)
        conf.conf_idle += 1;
      else
        conf.conf_idle = 0;
      if (conf.ctanum < 5 && conf.conf_lb == STRICT && !conf.force.has_lb &&
          !conf.policy_mlp)
        conf.conf_lb = CM;
    }
    if constexpr (std::is_same<G, device_graph_t<CSR, typename G::edge_t>>::value) {
      // fused levels skip the inspector, so the filter's per-vertex update
      // is applied to the queued vertices here (kernel_fusion.cuh:84-96)
      if (conf.conf_compensation && conf.conf_centric == VC &&
          conf.conf_fuse_inspect && !conf.conf_first_round &&
          as.fmt == Queue) {
        hipLaunchKernelGGL((gsw::k_compensation<G, F>),
                           dim3(gsw::grid_for(as.get_size_host())),
                           dim3(gsw::BLOCK), 0, gsw::stream(), as, g, f);
        H_ERR(hipGetLastError());
      }
    }
    const double e = gsw_now_ms();
    stat.level = g.level;
    stat.last_filter_time = e - s;
    stat.avg_filter_time =
        (stat.avg_filter_time * g.level + stat.last_filter_time) / (g.level + 1);
    fets.last_filter_time = stat.last_filter_time;
    fets.avg_filter_time = stat.avg_filter_time;
  }

  void run_filter(active_set_t &as, const config_t &conf) {
    const int64_t n = as.size;
    const int64_t nw = (n + 63) >> 6;
    const Status target = conf.conf_target;
    queue_t &q = as.queue;
    const int out_slot = (q.mode == Normal) ? (q.cursor ^ 1) : (q.cursor + 1);
    int32_t *out = (q.mode == Normal) ? q.buf[q.cursor ^ 1] : q.buf[0] + q.offset_out;
    int32_t *out_size = q.sizes + out_slot;
    const long long expect =
        (target == Active) ? as.known_active
        : (target == Inactive) ? as.known_inactive
                               : as.known_active + as.known_inactive;
    GSW_REQUIRE(expect <= (q.mode == Normal ? q.capacity : q.capacity - q.offset_out),
                gsw::E_OVERFLOW, "frontier does not fit the queue");
    gsw::set_one_async<int32_t>(out_size, 0);
    if (!conf.conf_sorted) {
      hipLaunchKernelGGL(gsw::k_filter_unsorted,
                         dim3(gsw::grid_for((nw + gsw::FILTER_WPT - 1) / gsw::FILTER_WPT,
                                            gsw::BLOCK, 1024)),
                         dim3(gsw::BLOCK), 0, gsw::stream(), as, target, n, out,
                         out_size);
    } else {
      const int nb = (int)MAX(1, MIN(1024, (nw + gsw::BLOCK - 1) / gsw::BLOCK));
      const int64_t wpb = (nw + nb - 1) / nb;
      as.ensure_block_count(nb);
      hipLaunchKernelGGL(gsw::k_filter_sorted_count, dim3(nb), dim3(gsw::BLOCK),
                         0, gsw::stream(), as, target, n, wpb, as.dg_block_count);
      hipLaunchKernelGGL(gsw::k_filter_sorted_write, dim3(nb), dim3(gsw::BLOCK),
                         0, gsw::stream(), as, target, n, wpb, as.dg_block_count,
                         out, out_size);
    }
    H_ERR(hipGetLastError());
    q.host_sizes[out_slot] = (int)expect;
  }

  // ------------------------------------------------------------- expand
  template <typename G, typename F>
  void expand(active_set_t &as, G &g, F &f, stat_t &stat, feature_t &fets,
              config_t &conf) {
    const double s = gsw_now_ms();
    // WM and CM scan the degrees of a wavefront's (64) or a workgroup's
    // (256) frontier vertices in int32; where the graph's degrees could
    // overflow that sum the level runs STRICT (int64 prefixes), forced
    // choices included
    if ((conf.conf_lb == WM || conf.conf_lb == CM) &&
        fets.max_deg * gsw::BLOCK > (double)INT32_MAX)
      conf.conf_lb = STRICT;
    level_record_t rec{};
    rec.level = g.level;
    rec.dir = conf.conf_dir;
    rec.fmt = as.fmt;
    rec.lb = conf.conf_lb;
    rec.fused = conf.conf_fuse_inspect;
    rec.sorted = conf.conf_sorted;
    rec.kernel = conf.conf_dir == Pull ? (!conf.mp_pull ? 1 : conf.pb_pull ? 2 : 0) : 0;
    rec.frontier = as.get_size_host();
    rec.active = as.known_active;
    rec.inactive = as.known_inactive;
    as.select_work_row((int)log.size());
    {
      // WM push: about 128 edges per wavefront at the frontier's mean degree
      // (the last inspected one on fused levels); measured against 64 edges
      // and against filling the capped grid first, this target keeps
      // long-range grids at fixed-64 speed and still halves ER levels
      const double d = fets.cur_avg_deg_active > 0 ? fets.cur_avg_deg_active
                                                   : fets.avg_deg;
      int vpw = gsw::WAVE;
      while (vpw > 8 && vpw * d > 128.0) vpw >>= 1;
      conf.wm_vpw = vpw;
      // early-exit TM pulls probe each target's first in-edge from the
      // per-vertex copy when a random in-edge is likely to end the walk:
      // the frontier holds >= 10% of the edges (on R-MAT, where the first
      // in-neighbour is a hub, far more walks end there)
      conf.pull_first = conf.conf_pruning && fets.push_workload_ratio >= 0.1;
    }
    int e0 = -1, e1 = -1;
    if (profile) {
      e0 = ev_next;
      H_ERR(hipEventRecord(next_event(), gsw::stream()));
    }
    dispatch(as, g, f, conf, stat);
    if (profile) {
      e1 = ev_next;
      H_ERR(hipEventRecord(next_event(), gsw::stream()));
    }
    log.push_back(rec);
    if (conf.verbose) {
      std::printf("level %d frontier %lld ", rec.level, rec.frontier);
      conf.show_hints(std::cout);
      std::cout << std::endl;
    }
    ev_pairs.push_back({e0, e1});

    if (conf.conf_fuse_inspect && conf.conf_centric == VC) {
      // the next queue's length: grid size for the next level, convergence
      const int out_slot = (as.queue.mode == Normal) ? (as.queue.cursor ^ 1)
                                                     : (as.queue.cursor + 1);
      const bool learned = conf.policy_mlp && stat.learned();
      int nq = as.read_qsize_device(out_slot);
      as.known_active = nq;
      as.check_overflow_cached();
      if (nq == 0) as.halt_host();
      // Learned selector on fused levels: a small next frontier stays fused
      // with the current class; a large one (where a pull or another format
      // may win) is handed to the next level's inspector, whose fused
      // inference picks the variant.
      if (learned) {
        if (conf.mlp_entered_fused) {  // the inspector chose fused for this level
          conf.mlp_fused_ref = rec.frontier;
          conf.mlp_entered_fused = false;
        }
        if (nq > 0 && (double)nq * gsw_mlp::INSPECT_FRACTION >= (double)as.size &&
            (double)nq >= gsw_mlp::INSPECT_GROWTH * (double)conf.mlp_fused_ref) {
          conf.conf_fuse_inspect = false;
          conf.conf_fusion = false;
          as.swap_queues();
        }
      }
    }
    const double e = gsw_now_ms();
    const int L = g.level > 0 ? g.level : 1;
    stat.last_expand_time = e - s;
    stat.last_time = stat.last_filter_time + stat.last_expand_time;
    stat.avg_time = (stat.avg_time * (L - 1) + stat.last_time) / L;
    stat.avg_expand_time = (stat.avg_expand_time * (L - 1) + stat.last_expand_time) / L;
    fets.last_expand_time = stat.last_expand_time;
    if (fets.pattern == Idem && fets.avg_expand_time > 0 &&
        fets.last_expand_time > 5 * fets.avg_expand_time &&
        conf.conf_fuse_inspect && !conf.force.has_fusion && !conf.policy_mlp) {
      conf.conf_fuse_inspect = false;
      as.swap_queues();
    }
    fets.avg_expand_time = stat.avg_expand_time;
    fets.last_time = stat.last_time;
    fets.avg_time = stat.avg_time;
    conf.conf_first_round = false;
  }

  template <typename G, typename F>
  void dispatch(active_set_t &as, G &g, F &f, config_t &conf, stat_t &stat);
};

// grid for a frontier of N items processed `per` items per workgroup
inline int expand_grid(int64_t N, int per) {
  int64_t g = (N + per - 1) / per;
  return (int)MAX(1, MIN((int64_t)gsw::MAX_GRID, g));
}

template <typename G, typename F>
void executor_t::dispatch(active_set_t &as, G &g, F &f, config_t &conf,
                          stat_t &stat) {
  using namespace gsw;
  hipStream_t st = stream();
  if constexpr (std::is_same<G, device_graph_t<COO, typename G::edge_t>>::value) {
    const int64_t N = as.get_size_host();
    hipLaunchKernelGGL((k_expand_ec<G, F>), dim3(grid_for(N)), dim3(BLOCK), 0,
                       st, as, g, f);
    H_ERR(hipGetLastError());
    return;
  } else {
    const int64_t N = as.get_size_host();
    if (N == 0) return;
    const bool push = conf.conf_dir == Push;
    const bool fused = conf.conf_fuse_inspect;
    LB lb = conf.conf_lb;
    // TWOD serves from-all pulls over a bitmap frontier of an unweighted
    // graph; anything else takes the edge-balanced STRICT path
    bool twod_ok = false;
    if constexpr (!G::has_edata)
      twod_ok = !push && conf.conf_fromall && as.fmt == Bitmap && g.twod;
    if (lb == ELB || (lb == TWOD && !twod_ok)) lb = STRICT;
    using MSG = typename pb_msg<F>::type;
    using WAt = typename F::wa_t;
    const bool use_pb = lb == TWOD && conf.pb_pull && g.pb;
    if (use_pb) {
      pb_ensure(g, PB_LDS_BYTES / (int)sizeof(MSG), PB_LDS_BYTES / (int)sizeof(WAt));
      pb_vals(*g.pb, sizeof(MSG));
    } else if (lb == TWOD) {
      // slice width from the bytes of source state one message reads
      constexpr int64_t src_bytes = sizeof(typename F::ra_t) > 4 ? sizeof(typename F::ra_t) : 4;
      twod_ensure(g, src_bytes);
    }
    // Scratch must exist before any kernel body copies `as` by value.
    if (lb == TWC && push) as.bins.ensure(as.size);
    if (lb == STRICT) as.workset.ensure(as.size);
    GSW_REQUIRE(!(lb == TWC && push) || (as.bins.sizes && as.bins.bin[2]),
                E_INTERNAL_CHECK, "TWC bins missing");
    GSW_REQUIRE(lb != STRICT || (as.workset.dg_scan && as.workset.dg_bsum),
                E_INTERNAL_CHECK, "STRICT workset missing");
    if (push) {
      auto launch = [&](auto body) {
        using Bd = decltype(body);
        switch (lb) {
          case TM:
            hipLaunchKernelGGL(k_push_tm<Bd>, dim3(expand_grid(N, BLOCK)),
                               dim3(BLOCK), 0, st, body);
            break;
          case WM:
            hipLaunchKernelGGL(k_push_wm<Bd>,
                               dim3(expand_grid(N * (WAVE / conf.wm_vpw), BLOCK)),
                               dim3(BLOCK), 0, st, body, conf.wm_vpw);
            break;
          case CM:
            hipLaunchKernelGGL(k_push_cm<Bd>, dim3(expand_grid(N, BLOCK)),
                               dim3(BLOCK), 0, st, body);
            break;
          case TWC: {
            fill_async<int32_t>(as.bins.sizes, 0, 4);
            hipLaunchKernelGGL(k_twc_bin<G>, dim3(expand_grid(N, BLOCK)),
                               dim3(BLOCK), 0, st, as, g, conf.conf_target,
                               false, conf.twc_small, conf.twc_large);
            hipLaunchKernelGGL(k_push_twc<Bd>, dim3(expand_grid(N, BLOCK)),
                               dim3(BLOCK), 0, st, body);
            break;
          }
          default: {  // STRICT
            const int64_t *scan = nullptr;  // bitmap: implicit rows[v] + v
            if (as.fmt == Queue) {
              hipLaunchKernelGGL(k_strict_prepare<G>, dim3(grid_for(N)),
                                 dim3(BLOCK), 0, st, as, g, conf.conf_target,
                                 false, as.workset.dg_scan);
              scan_exclusive_inplace(as.workset.dg_scan, N, as.workset.dg_bsum);
              scan = as.workset.dg_scan;
            }
            hipLaunchKernelGGL(k_push_strict<Bd>, dim3(MAX_GRID), dim3(BLOCK),
                               0, st, body, scan);
            break;
          }
        }
        H_ERR(hipGetLastError());
      };
      if (fused)
        launch(push_body<G, F, true>{as, g, f, conf});
      else
        launch(push_body<G, F, false>{as, g, f, conf});
    } else {
      using Bd = pull_body<G, F>;
      // (one group only: with several, a later group's phase 1 would read
      // ranks an earlier group's filter had already updated)
      conf.pb_fuse_filter = use_pb && conf.counts_known && g.pb->ng == 1;
      Bd body{as, g, f, conf};
      switch (lb) {
        case TM:
          if constexpr (is_level_sync<F>::value) {
            if (as.fmt == Bitmap && !conf.conf_fromall && !conf.conf_toall) {
              // this level's pull also inspects the next one
              gsw_mlp::sel_dev_t d;
              if (conf.policy_mlp && stat.mlp.w) {
                d = stat.mlp;
                d.level = (float)g.level;
              }
              hipLaunchKernelGGL(k_pull_tm_inspect<Bd>,
                                 dim3(grid_for((g.nvertexs + 63) / 64 * 64)),
                                 dim3(BLOCK), 0, st, body,
                                 as.bitmap.visited.dg_bits, stat.slots());
              hipLaunchKernelGGL(k_stat_finish, dim3(1), dim3(BLOCK), 0, st, d,
                                 stat.slots(), stat.dg_runtime_info);
              H_ERR(hipGetLastError());
              // the next frontier becomes `active`; the old active words
              // (vertices reached earlier) become the `visited` marks
              std::swap(as.bitmap.active.dg_bits, as.bitmap.visited.dg_bits);
              conf.conf_pre_inspected = true;
              break;
            }
          }
          hipLaunchKernelGGL(k_pull_tm<Bd>, dim3(expand_grid(N, BLOCK)),
                             dim3(BLOCK), 0, st, body);
          break;
        case WM:
        case TWC:
          hipLaunchKernelGGL(k_pull_wm<Bd>, dim3(expand_grid(N, BLOCK)),
                             dim3(BLOCK), 0, st, body);
          break;
        case CM:
          hipLaunchKernelGGL(k_pull_cm<Bd>, dim3(expand_grid(N, BLOCK)),
                             dim3(BLOCK), 0, st, body);
          break;
        case TWOD: {
          if (use_pb) {
            // propagation blocking: messages to slots, then slots to targets
            // group by group (consecutive bins): a group's phase 2 follows
            // its own phase 1
            const pb_t &pb = *g.pb;
            for (int64_t k = 0; k < pb.ng; ++k) {
              pb_t q = pb;
              q.p1 = pb.p1 + 4 * pb.g1h[k];
              q.p2 = pb.p2 + 4 * pb.g2h[k];
              const int64_t n1 = pb.g1h[k + 1] - pb.g1h[k];
              const int64_t n2 = pb.g2h[k + 1] - pb.g2h[k];
              if (n1 > 0)
                hipLaunchKernelGGL((k_pb_scatter<Bd, MSG>), dim3(n1), dim3(PB_BLOCK),
                                   0, st, body, q);
              if (n2 > 0)
                hipLaunchKernelGGL((k_pb_gather<Bd, MSG>), dim3(n2), dim3(PB_BLOCK), 0,
                                   st, body, q);
            }
            if (conf.pb_fuse_filter) {
              // targets of bins split over workgroups: filtered once all
              // parts are added
              if (pb.nsplit > 0)
                hipLaunchKernelGGL((k_pb_filter_split<Bd>),
                                   dim3(grid_for(pb.nsplit)), dim3(BLOCK), 0, st, body,
                                   (const int32_t *)pb.split, pb.nsplit);
              conf.filter_done = true;
            }
            break;
          }
          // the slices one after another (each launch owns its targets)
          const twod_t &tw = *g.twod;
          for (int p = 0; p < tw.P; ++p) {
            const twod_slice_t &sl = tw.s[p];
            if (sl.ne == 0) continue;
            if (conf.mp_pull) {
              int grid = 0;
              for (int c = 0; c <= TWOD_BINS; ++c) grid += sl.blocks[c];
              hipLaunchKernelGGL(k_pull_bins<Bd>, dim3(grid), dim3(BLOCK), 0, st,
                                 body, sl);
            } else {
              const int64_t tiles = (sl.ne + sl.nr + SCHUNK - 1) / SCHUNK;
              const int grid = (int)MIN((int64_t)MAX_GRID, tiles);
              hipLaunchKernelGGL(k_pull_twod<Bd>, dim3(grid), dim3(BLOCK), 0, st,
                                 body, sl);
            }
          }
          break;
        }
        default: {  // STRICT
          const int64_t *scan = nullptr;  // bitmap: implicit rows[v] + v
          if (conf.mp_pull && conf.conf_fromall && !conf.conf_pruning &&
              as.fmt == Bitmap) {
            // every in-edge contributes: the merge-path walk is cheaper
            hipLaunchKernelGGL(k_pull_mp<Bd>, dim3(MAX_GRID), dim3(BLOCK), 0,
                               st, body);
            break;
          }
          if (as.fmt == Queue) {
            hipLaunchKernelGGL(k_strict_prepare<G>, dim3(grid_for(N)),
                               dim3(BLOCK), 0, st, as, g, conf.conf_target,
                               true, as.workset.dg_scan);
            scan_exclusive_inplace(as.workset.dg_scan, N, as.workset.dg_bsum);
            scan = as.workset.dg_scan;
          }
          hipLaunchKernelGGL(k_pull_strict<Bd>, dim3(MAX_GRID), dim3(BLOCK), 0,
                             st, body, scan);
          break;
        }
      }
      H_ERR(hipGetLastError());
    }
  }
}
