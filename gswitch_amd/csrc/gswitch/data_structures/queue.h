// Frontier queues.
//
// queue_t keeps the reference's two modes (src/data_structures/queue.cuh:
// 86-188): Normal = ping-pong input/output queues swapped every level;
// Cached = every level's queue appended to one history array (used by BC's
// backward "traceback" replay). Differences: pushes are staged per
// wavefront in LDS (one atomic per flushed block of ids) and bounds-checked against the output capacity; an
// overflow sets a device flag the host turns into an error instead of
// writing past the allocation (queue.cuh:65-68 has no check).
#pragma once
#include <cstdlib>

#include "../utils/common.h"
#include "../utils/wave.h"

struct queue_t {
  int32_t *buf[2] = {nullptr, nullptr};
  int32_t *sizes = nullptr;     // device: Normal [2]; Cached [max_levels]
  int32_t *overflow = nullptr;  // device flag
  int64_t capacity = 0;
  int max_levels = 0;
  int cursor = 0;
  int64_t offset_in = 0, offset_out = 0;
  QueueMode mode = Normal;
  bool traceback = false;

  int64_t build(int64_t cap, QueueMode qmode, int32_t *overflow_flag,
                int levels = 4096) {
    mode = qmode;
    capacity = cap;
    overflow = overflow_flag;
    max_levels = (mode == Cached) ? levels : 2;
    H_ERR(hipMalloc((void **)&buf[0], sizeof(int32_t) * (capacity + 1)));
    if (mode == Normal)
      H_ERR(hipMalloc((void **)&buf[1], sizeof(int32_t) * (capacity + 1)));
    H_ERR(hipMalloc((void **)&sizes, sizeof(int32_t) * (max_levels + 2)));
    H_ERR(hipMemset(sizes, 0, sizeof(int32_t) * (max_levels + 2)));
    host_sizes = (int *)calloc(max_levels + 2, sizeof(int));
    cursor = 0;
    offset_in = offset_out = 0;
    return sizeof(int32_t) * capacity * (mode == Normal ? 2 : 1);
  }
  void release() {
    hipFree(buf[0]);
    if (buf[1]) hipFree(buf[1]);
    hipFree(sizes);
    free(host_sizes);
    host_sizes = nullptr;
    buf[0] = buf[1] = nullptr;
    sizes = nullptr;
  }

  // ---- device view ----
  __device__ __forceinline__ int get_qsize() const { return sizes[cursor]; }
  __device__ __forceinline__ int fetch(int idx) const {
    return (mode == Normal) ? buf[cursor][idx] : buf[0][offset_in + idx];
  }
  __device__ __forceinline__ int32_t *output_base() const {
    return (mode == Normal) ? buf[cursor ^ 1] : buf[0] + offset_out;
  }
  __device__ __forceinline__ int32_t *output_size() const {
    return (mode == Normal) ? &sizes[cursor ^ 1] : &sizes[cursor + 1];
  }
  __device__ __forceinline__ int64_t output_capacity() const {
    return (mode == Normal) ? capacity : capacity - offset_out;
  }
  // Append u to the output queue, staged per wavefront in LDS (no-op during
  // traceback): all 64 lanes call it; the entries reach the queue when the
  // stage fills and at flush_stage().
  static constexpr int STAGE = 512;  // entries per wavefront
  using stage_t = gsw::wave_stage_t<STAGE>;
  __device__ __forceinline__ void push(bool want, int u, stage_t &st) {
    if (mode == Cached && traceback) return;
    st.put(want, u, output_size(), output_base(), output_capacity(), overflow);
  }
  __device__ __forceinline__ void flush_stage(stage_t &st) {
    if (mode == Cached && traceback) return;
    st.flush(output_size(), output_base(), output_capacity(), overflow);
  }
  // the workgroup's final flush (all threads call it; s_cnt: waves + 1 ints)
  __device__ __forceinline__ void flush_stage_block(stage_t &st, int *s_cnt) {
    if (mode == Cached && traceback) return;
    st.flush_block(output_size(), output_base(), output_capacity(), overflow, s_cnt);
  }

  // ---- host side (a plain array: queue_t is copied byte-wise to kernels) --
  int *host_sizes = nullptr;  // sizes known on the host, per slot
};
