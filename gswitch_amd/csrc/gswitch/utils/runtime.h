// Process-wide runtime pieces: the engine's HIP stream, fill kernels and
// small launch helpers. All engine work is ordered on one stream per device;
// host reads of device scalars go through pinned memory + one stream sync.
#pragma once
#include "common.h"

namespace gsw {

inline hipStream_t &stream_slot() {
  static thread_local hipStream_t s = nullptr;
  return s;
}

// The stream every engine launch uses (created on first use, default
// priority, blocking w.r.t. the null stream so plain hipMemcpy orders with it).
inline hipStream_t stream() {
  hipStream_t &s = stream_slot();
  if (!s) H_ERR(hipStreamCreate(&s));
  return s;
}

template <typename T>
__global__ void k_fill(T *__restrict__ p, T val, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = val;
}

inline int grid_for(int64_t work, int block = BLOCK, int max_grid = MAX_GRID) {
  int64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_grid) g = max_grid;
  return (int)g;
}

// Device-side fill (a kernel, not hipMemsetAsync, so it can live in a
// captured graph).
template <typename T> inline void fill_async(T *p, T val, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n)), dim3(BLOCK), 0, stream(),
                     p, val, n);
  H_ERR(hipGetLastError());
}

template <typename T> inline void set_one_async(T *p, T val) {
  hipLaunchKernelGGL(k_fill<T>, dim3(1), dim3(64), 0, stream(), p, val,
                     (int64_t)1);
  H_ERR(hipGetLastError());
}

inline void sync() { H_ERR(hipStreamSynchronize(stream())); }

// Pinned host staging for small device->host reads.
struct pinned_t {
  void *p = nullptr;
  size_t bytes = 0;
  void *get(size_t b) {
    if (b > bytes) {
      if (p) hipHostFree(p);
      H_ERR(hipHostMalloc(&p, b));
      bytes = b;
    }
    return p;
  }
  ~pinned_t() {
    if (p) hipHostFree(p);
  }
};

}  // namespace gsw
