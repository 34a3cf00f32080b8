// The frontier ("active set").
//
// Mirrors active_set_t (src/data_structures/active_set.cuh:14-102): three
// bitmaps, the queue pair, the STRICT workset and the convergence state,
// with the same host API (build / reset / init / finish / get_size_host /
// halt_host) and device API (get_size / fetch / halt_device). Differences:
//  * the halt flag and the overflow flag live in device memory (a control
//    block), not in mapped host memory, so level loops can be captured;
//  * the 268 MB always-allocated `bins` scratch of the reference
//    (queue.cuh:10-25) is replaced by three degree bins of n entries used by
//    the TWC expand, allocated on first use;
//  * STRICT's workset holds int64 prefix offsets.
#pragma once
#include <algorithm>

#include "../abstraction/config.h"
#include "../utils/runtime.h"
#include "bitmap.h"
#include "queue.h"

// STRICT partition scratch (workset_t, src/data_structures/workset.cuh:8-45)
struct workset_t {
  int64_t *dg_scan = nullptr;  // [cap + 1] exclusive prefix of (deg + 1)
  int64_t *dg_bsum = nullptr;  // [max_blocks + 1] device-scan block sums
  int64_t cap = 0;
  int max_blocks = 0;

  void ensure(int64_t n) {
    if (dg_scan) return;
    cap = n;
    max_blocks = (int)((n + 2047) / 2048) + 1;
    H_ERR(hipMalloc((void **)&dg_scan, sizeof(int64_t) * (cap + 2)));
    H_ERR(hipMalloc((void **)&dg_bsum, sizeof(int64_t) * (max_blocks + 2)));
  }
  void release() {
    if (dg_scan) hipFree(dg_scan);
    if (dg_bsum) hipFree(dg_bsum);
    dg_scan = nullptr;
    dg_bsum = nullptr;
  }
};

// TWC degree bins: small (thread), medium (wavefront), large (workgroup).
struct bins_t {
  int32_t *bin[3] = {nullptr, nullptr, nullptr};
  int32_t *sizes = nullptr;  // device [3]
  int64_t cap = 0;

  void ensure(int64_t n) {
    if (sizes) return;
    cap = n;
    for (int i = 0; i < 3; ++i)
      H_ERR(hipMalloc((void **)&bin[i], sizeof(int32_t) * (cap + 1)));
    H_ERR(hipMalloc((void **)&sizes, sizeof(int32_t) * 4));
  }
  void release() {
    for (int i = 0; i < 3; ++i)
      if (bin[i]) hipFree(bin[i]);
    if (sizes) hipFree(sizes);
    bin[0] = bin[1] = bin[2] = nullptr;
    sizes = nullptr;
  }
};

// mark roots[0..k) active (and not inactive) in both frontier bitmaps
__global__ void __launch_bounds__(256)
    k_set_roots(uint64_t *active, uint64_t *inactive, const int32_t *roots,
                int64_t k) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int v = roots[i];
    atomicOr((unsigned long long *)&active[v >> 6], 1ull << (v & 63));
    atomicOr((unsigned long long *)&inactive[v >> 6], 1ull << (v & 63));
  }
}

enum : int { CTRL_HALT = 0, CTRL_OVERFLOW = 1, CTRL_FLAG = 2, CTRL_N = 8 };
// Per-expand work counters (edges scanned, vertices expanded): one record
// row of WORK_N counters per expand, WORK_ROWS rows per traversal, kept as
// WORK_SLOTS partial copies laid out slot-major ([slot][row][counter]) so
// that the slots of one row sit on different cache lines. k_work_fold sums
// the slots into slot 0 when the records are read.
enum : int { WORK_EDGES = 0, WORK_VERTS = 1, WORK_N = 4 };
constexpr int WORK_ROWS = 16384;  // per-expand counter rows kept per traversal
constexpr int WORK_SLOTS = 16;
constexpr int64_t WORK_SLOT_STRIDE = (int64_t)WORK_ROWS * WORK_N;
constexpr int64_t WORK_TOTAL = WORK_SLOT_STRIDE * WORK_SLOTS;

// A traversal's reset in one launch: zeroes the three frontier bitmaps
// (n_bits words each), the control block, the work counters and the first
// nq queue lengths; n_all is the largest of these counts.
__global__ void __launch_bounds__(256)
    k_reset_frontier(uint64_t *visited, uint64_t *active, uint64_t *inactive,
                     int64_t n_bits, int32_t *ctrl, unsigned long long *work,
                     int32_t *qsizes, int64_t nq, int64_t n_all) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_all;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < n_bits) {
      visited[i] = 0;
      active[i] = 0;
      inactive[i] = 0;
    }
    if (i < CTRL_N) ctrl[i] = 0;
    if (i < WORK_TOTAL) work[i] = 0;
    if (i < nq) qsizes[i] = 0;
  }
}

// one root active (and not inactive) on cleared frontier bitmaps
__global__ void __launch_bounds__(64)
    k_set_root(uint64_t *active, uint64_t *inactive, int64_t root) {
  if (threadIdx.x == 0) {
    active[root >> 6] = 1ull << (root & 63);
    inactive[root >> 6] = 1ull << (root & 63);
  }
}

// zero record row `row` in every slot
__global__ void __launch_bounds__(64)
    k_work_clear_row(unsigned long long *row) {
  const int s = threadIdx.x / WORK_N, k = threadIdx.x % WORK_N;
  if (s < WORK_SLOTS) row[s * WORK_SLOT_STRIDE + k] = 0;
}

// slot 0 of rows [0, nrows) := the sum over the slots; the other slots are
// cleared, so folding again changes nothing
__global__ void __launch_bounds__(256)
    k_work_fold(unsigned long long *base, int64_t nrows) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows * WORK_N) return;
  unsigned long long t = base[i];
  for (int s = 1; s < WORK_SLOTS; ++s) {
    t += base[s * WORK_SLOT_STRIDE + i];
    base[s * WORK_SLOT_STRIDE + i] = 0;
  }
  base[i] = t;
}

struct active_set_t {
  int64_t size = 0;  // number of vertices (VC) or edges (EC)
  ASFmt fmt = Queue;
  int64_t size_cached = 0;
  bitmap_t bitmap;
  queue_t queue;
  workset_t workset;
  bins_t bins;
  int32_t *dg_ctrl = nullptr;                 // device control block
  unsigned long long *dg_work = nullptr;      // counters of the current expand
  unsigned long long *dg_work_base = nullptr; // [WORK_SLOTS][WORK_ROWS][WORK_N]
  int32_t *dg_block_count = nullptr;          // sorted-filter scratch [1025]
  int32_t *h_ctrl = nullptr;                  // pinned mirror of dg_ctrl
  // host-side counts of the current frontier (from the inspector's
  // counters, or the root): the filter sizes its output queue with them
  // without a device round trip
  long long known_active = 0, known_inactive = 0;
  bool converged = false;                     // host view of the halt flag
  bool window_moved = false;                  // finish() advanced the window

  int64_t build(int64_t n, QueueMode qmode) {
    size = n;
    int64_t bytes = 0;
    H_ERR(hipMalloc((void **)&dg_ctrl, sizeof(int32_t) * CTRL_N));
    H_ERR(hipMalloc((void **)&dg_work_base, sizeof(unsigned long long) * WORK_TOTAL));
    dg_work = dg_work_base;
    H_ERR(hipHostMalloc((void **)&h_ctrl, sizeof(int32_t) * CTRL_N));
    bytes += queue.build(n, qmode, dg_ctrl + CTRL_OVERFLOW);
    bytes += bitmap.build(n);
    reset();
    return bytes;
  }

  void release() {
    queue.release();
    bitmap.release();
    workset.release();
    bins.release();
    hipFree(dg_ctrl);
    hipFree(dg_work_base);
    if (dg_block_count) hipFree(dg_block_count);
    hipHostFree(h_ctrl);
    dg_ctrl = nullptr;
    dg_work = dg_work_base = nullptr;
    dg_block_count = nullptr;
    h_ctrl = nullptr;
  }

  // Clear frontier state for a new traversal (traceback keeps the Cached
  // history and its cursor, as queue_t::reset does in the reference).
  void reset(bool traceback = false) {
    queue.traceback = traceback;
    int64_t nq = 0;  // queue lengths to clear
    if (queue.mode == Normal) {
      nq = queue.max_levels + 2;
      queue.host_sizes[0] = queue.host_sizes[1] = 0;
    }
    if (!traceback) {
      queue.cursor = 0;
      queue.offset_in = queue.offset_out = 0;
      if (queue.mode == Cached) {
        nq = queue.max_levels + 2;
        for (int i = 0; i < queue.max_levels + 2; ++i) queue.host_sizes[i] = 0;
      }
    }
    const int64_t nb = bitmap.visited.n_words + 1;
    GSW_REQUIRE(bitmap.active.n_words + 1 == nb && bitmap.inactive.n_words + 1 == nb,
                gsw::E_INTERNAL_CHECK, "frontier bitmaps of different sizes");
    const int64_t n_all = std::max(std::max(nb, nq), std::max((int64_t)CTRL_N, WORK_TOTAL));
    hipLaunchKernelGGL(k_reset_frontier, dim3(gsw::grid_for(n_all, 256)), dim3(256),
                       0, gsw::stream(), bitmap.visited.dg_bits,
                       bitmap.active.dg_bits, bitmap.inactive.dg_bits, nb, dg_ctrl,
                       dg_work_base, queue.sizes, nq, n_all);
    H_ERR(hipGetLastError());
    dg_work = dg_work_base;
    converged = false;
    window_moved = false;
  }

  // root >= 0: only `root` is active; ALL_ACTIVE: every vertex; ALL_INACTIVE:
  // none (bitmap_t::init, src/data_structures/bitmap.cuh:133-144).
  // after_reset: called right after reset(), which left both bitmaps zero.
  void init(int64_t root = ALL_ACTIVE, bool after_reset = false) {
    const int64_t nw = bitmap.active.n_words;
    if (root >= 0) {
      GSW_REQUIRE(root < size, gsw::E_INVALID, "root out of range");
      if (!after_reset) {
        gsw::fill_async<uint64_t>(bitmap.active.dg_bits, 0, nw + 1);
        gsw::fill_async<uint64_t>(bitmap.inactive.dg_bits, 0, nw + 1);
      }
      hipLaunchKernelGGL(k_set_root, dim3(1), dim3(64), 0, gsw::stream(),
                         bitmap.active.dg_bits, bitmap.inactive.dg_bits, root);
      H_ERR(hipGetLastError());
    } else if (root == ALL_ACTIVE) {
      gsw::fill_async<uint64_t>(bitmap.active.dg_bits, ~0ull, nw + 1);
      gsw::fill_async<uint64_t>(bitmap.inactive.dg_bits, ~0ull, nw + 1);
    } else if (!after_reset) {
      gsw::fill_async<uint64_t>(bitmap.active.dg_bits, 0, nw + 1);
      gsw::fill_async<uint64_t>(bitmap.inactive.dg_bits, 0, nw + 1);
    }
  }

  // Several roots at once (a batch of disjoint instances, or multi-source
  // traversals): `d_roots` holds k distinct vertex ids on the device.
  void init_roots(const int32_t *d_roots, int64_t k, bool after_reset = false) {
    const int64_t nw = bitmap.active.n_words;
    if (!after_reset) {
      gsw::fill_async<uint64_t>(bitmap.active.dg_bits, 0, nw + 1);
      gsw::fill_async<uint64_t>(bitmap.inactive.dg_bits, 0, nw + 1);
    }
    if (k > 0) {
      hipLaunchKernelGGL(k_set_roots, dim3(gsw::grid_for(k)), dim3(256), 0,
                         gsw::stream(), bitmap.active.dg_bits,
                         bitmap.inactive.dg_bits, d_roots, k);
      H_ERR(hipGetLastError());
    }
  }

  bool is_converged() const { return converged; }
  void halt_host() { converged = true; }
  void set_fmt(ASFmt f) { fmt = f; }

  // Host view of the current input size: queue length or n for bitmaps.
  int64_t get_size_host() {
    size_cached = (fmt == Queue) ? queue.host_sizes[queue.cursor] : size;
    return size_cached;
  }
  int64_t get_size_host_cached() const { return size_cached; }

  // Read the device queue length of `slot` together with the control block
  // (one stream sync); raises on a queue overflow.
  int read_qsize_device(int slot) {
    int32_t *h = h_ctrl;
    H_ERR(hipMemcpyAsync(h, dg_ctrl, sizeof(int32_t) * CTRL_FLAG,
                         hipMemcpyDeviceToHost, gsw::stream()));
    H_ERR(hipMemcpyAsync(h + CTRL_FLAG, queue.sizes + slot, sizeof(int32_t),
                         hipMemcpyDeviceToHost, gsw::stream()));
    gsw::sync();
    queue.host_sizes[slot] = h[CTRL_FLAG];
    GSW_REQUIRE(h[CTRL_OVERFLOW] == 0, gsw::E_OVERFLOW, "frontier queue overflow");
    return h[CTRL_FLAG];
  }

  void check_overflow() {
    H_ERR(hipMemcpyAsync(h_ctrl, dg_ctrl, sizeof(int32_t) * CTRL_FLAG,
                         hipMemcpyDeviceToHost, gsw::stream()));
    gsw::sync();
    GSW_REQUIRE(h_ctrl[CTRL_OVERFLOW] == 0, gsw::E_OVERFLOW,
                "frontier queue overflow");
  }
  // overflow state as of the last read_qsize_device (no extra sync)
  void check_overflow_cached() const {
    GSW_REQUIRE(h_ctrl[CTRL_OVERFLOW] == 0, gsw::E_OVERFLOW,
                "frontier queue overflow");
  }

  // route the next expand's work counters to record row `i`
  // true once after finish() advanced the stepping window
  bool take_window_moved() {
    const bool m = window_moved;
    window_moved = false;
    return m;
  }

  void select_work_row(int i) {
    dg_work = dg_work_base + (int64_t)(i < WORK_ROWS ? i : WORK_ROWS - 1) * WORK_N;
    // past the last row every expand reuses it: start it from zero
    if (i >= WORK_ROWS - 1) {
      hipLaunchKernelGGL(k_work_clear_row, dim3(1), dim3(64), 0, gsw::stream(),
                         dg_work);
      H_ERR(hipGetLastError());
    }
  }

  void ensure_block_count(int nb) {
    if (!dg_block_count)
      H_ERR(hipMalloc((void **)&dg_block_count, sizeof(int32_t) * 1025));
    (void)nb;
  }

  // Advance the queues after the current input has been consumed
  // (queue_t::swap, src/data_structures/queue.cuh:113-127).
  void swap_queues() {
    if (queue.mode == Normal) {
      gsw::set_one_async<int32_t>(queue.sizes + queue.cursor, 0);
      queue.host_sizes[queue.cursor] = 0;
      queue.cursor ^= 1;
    } else if (queue.traceback) {
      GSW_REQUIRE(queue.cursor > 0, gsw::E_INTERNAL_CHECK,
                  "history queue traced back past its first level");
      queue.cursor--;
      queue.offset_in -= queue.host_sizes[queue.cursor];
    } else {
      // the next expand writes the size slot after the new cursor; both
      // size arrays hold max_levels + 2 slots
      GSW_REQUIRE(queue.cursor < queue.max_levels, gsw::E_OVERFLOW,
                  "history queue holds no more levels");
      queue.cursor++;
      queue.offset_in = queue.offset_out;
      queue.offset_out += queue.host_sizes[queue.cursor];
    }
  }

  template <typename G, typename F> bool finish(G &g, F &f, config_t &conf);

  // ---- device view ----
  __device__ __forceinline__ int64_t get_size() const {
    return fmt == Queue ? (int64_t)queue.get_qsize() : size;
  }
  // Frontier element idx, or -1 if (bitmap format) it is not of status
  // `wanted` (ASProxy / fetch_bitmap, active_set.cuh:58-80).
  __device__ __forceinline__ int fetch(int64_t idx, Status wanted) const {
    if (fmt == Queue) return queue.fetch((int)idx);
    int v = (int)idx;
    bool ok = (wanted == Active)     ? bitmap.is_active(v)
              : (wanted == Inactive) ? bitmap.is_inactive(v)
                                     : (bitmap.is_active(v) || bitmap.is_inactive(v));
    return ok ? v : -1;
  }
  __device__ __forceinline__ void halt_device() { dg_ctrl[CTRL_HALT] = 1; }
};

inline active_set_t build_active_set(int64_t size, const config_t &conf) {
  active_set_t as;
  as.build(size, conf.conf_qmode);
  return as;
}
