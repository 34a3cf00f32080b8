// Edge-centric expand, pointer jumping, and the small functor passes the
// level loop needs around fusion and the SSSP window.
//  k_expand_ec          __expand_EC (src/kernel_libs/expand_EC.cuh:11-29)
//  k_pointer_jumping    pointer_jumping (src/utils/intrinsics.cuh:328-349)
//  k_filter_all         __window_delay_update (active_set.cuh:167-174)
//  k_exit_any           __exit (src/data_structures/notification.cuh:9-31)
//  k_compensation       __compensation_for_queue (kernel_fusion.cuh:54-65)
#pragma once
#include "../data_structures/active_set.h"
#include "../data_structures/functor.h"
#include "../utils/wave.h"

namespace gsw {

template <typename G, typename F>
__global__ void __launch_bounds__(BLOCK) k_expand_ec(active_set_t as, G g, F f) {
  const int64_t N = as.get_size();
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < N;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int e = as.fetch(idx, Active);
    if (e < 0) continue;
    const packed_t vv = g.dg_coo[e];
    f.update((int)(vv >> 32), (int)(vv & 0xffffffffll),
             (typename G::edge_t *)nullptr);
  }
}

// lab[v] <- lab[lab[v]]; *changed = 1 if any label moved
__global__ void __launch_bounds__(BLOCK)
    k_pointer_jumping(int *lab, int64_t n, int *changed) {
  bool moved = false;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int y = lab[v];
    const int x = lab[y];
    if (x != y) {
      lab[v] = x;
      moved = true;
    }
  }
  if (ballot(moved) != 0 && lane_id() == 0) *changed = 1;
}

template <typename G, typename F>
__global__ void __launch_bounds__(BLOCK) k_filter_all(G g, F f) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       v < g.nvertexs; v += (int64_t)gridDim.x * blockDim.x)
    f.filter((int)v, g);
}

template <typename G, typename F>
__global__ void __launch_bounds__(BLOCK) k_exit_any(G g, F f, int *flag) {
  bool any = false;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       v < g.nvertexs; v += (int64_t)gridDim.x * blockDim.x)
    any |= f.exit((int)v, g);
  if (ballot(any) != 0 && lane_id() == 0) *flag = 1;
}

template <typename G, typename F>
__global__ void __launch_bounds__(BLOCK)
    k_compensation(active_set_t as, G g, F f) {
  const int64_t N = as.queue.get_qsize();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    f.filter(as.queue.fetch((int)i), g);
}

}  // namespace gsw
