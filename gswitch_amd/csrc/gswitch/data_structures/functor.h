// The user operator API.
//
// Same shape as the reference (src/data_structures/functor.cuh:12-98):
// vdata_t<WA, RA> owns the per-vertex write/read attributes, Functor<VC, ...>
// exposes filter / emit / cond / comp / compAtomic / exit and Functor<EC, ...>
// exposes filter / update. The base methods are non-virtual defaults: the
// engine's kernels are templated on the concrete functor, so a derived
// struct's methods hide them and inline into the traversal loops (no vtable
// is ever copied to the device).
#pragma once
#include <cstdlib>
#include <type_traits>

#include "../utils/common.h"
#include "graph.h"
#include "window.h"

template <typename WA, typename RA> struct vdata_t {
  static constexpr bool has_wa = !std::is_same<WA, Empty>::value;
  static constexpr bool has_ra = !std::is_same<RA, Empty>::value;

  void build(int64_t n) {
    size = n;
    if (has_wa) {
      H_ERR(hipMalloc((void **)&dg_wa, sizeof(WA) * (size + 1)));
      h_wa = (WA *)malloc(sizeof(WA) * (size + 1));
    }
    if (has_ra) {
      H_ERR(hipMalloc((void **)&dg_ra, sizeof(RA) * (size + 1)));
      h_ra = (RA *)malloc(sizeof(RA) * (size + 1));
    }
  }

  void release() {
    if (has_wa) {
      hipFree(dg_wa);
      free(h_wa);
    }
    if (has_ra) {
      hipFree(dg_ra);
      free(h_ra);
    }
    dg_wa = nullptr;
    dg_ra = nullptr;
    h_wa = nullptr;
    h_ra = nullptr;
  }

  template <typename X> void init_wa(X init) {
    for (int64_t i = 0; i < size; ++i) h_wa[i] = init((int)i);
    H_ERR(hipMemcpy(dg_wa, h_wa, sizeof(WA) * size,
                    hipMemcpyHostToDevice));
  }
  template <typename X> void init_ra(X init) {
    for (int64_t i = 0; i < size; ++i) h_ra[i] = init((int)i);
    H_ERR(hipMemcpy(dg_ra, h_ra, sizeof(RA) * size,
                    hipMemcpyHostToDevice));
  }
  void set_zero(WA z) { zero = z; }
  void sync_wa() {
    H_ERR(hipMemcpy(h_wa, dg_wa, sizeof(WA) * size,
                    hipMemcpyDeviceToHost));
  }
  void sync_ra() {
    H_ERR(hipMemcpy(h_ra, dg_ra, sizeof(RA) * size,
                    hipMemcpyDeviceToHost));
  }

  __device__ __forceinline__ WA *fetch_wa(int v) const { return dg_wa + v; }
  __device__ __forceinline__ RA *fetch_ra(int v) const { return dg_ra + v; }

  WA zero{};
  WA *dg_wa = nullptr;
  RA *dg_ra = nullptr;
  WA *h_wa = nullptr;  // host mirrors (plain arrays: the functor is copied
  RA *h_ra = nullptr;  // byte-wise into every kernel)
  int64_t size = 0;
  window_t window;
};

template <Centric C, typename WA, typename RA, typename E> struct Functor {};

// A functor is level-synchronous when a vertex updated during a level is
// exactly the next level's frontier and every other vertex keeps its status
// (BFS-like traversals). Such a functor may declare
//   static constexpr bool level_sync = true;
// and the engine then folds the next level's inspection into TM pulls.
template <typename F, typename = void> struct is_level_sync : std::false_type {};
template <typename F>
struct is_level_sync<F, std::void_t<decltype(F::level_sync)>>
    : std::integral_constant<bool, F::level_sync> {};

template <typename WA, typename RA, typename E>
struct Functor<EC, WA, RA, E> {
  typedef WA wa_t;
  typedef RA ra_t;
  typedef E e_t;
  using G = device_graph_t<COO, E>;
  vdata_t<WA, RA> data;

  __device__ __forceinline__ WA *wa_of(int v) const { return data.fetch_wa(v); }
  __device__ __forceinline__ RA *ra_of(int v) const { return data.fetch_ra(v); }

  __device__ Status filter(int v, int u, E *e) { return Inactive; }
  __device__ void update(int v, int u, E *e) {}
};

template <typename WA, typename RA, typename E>
struct Functor<VC, WA, RA, E> {
  typedef WA wa_t;
  typedef RA ra_t;
  typedef E e_t;
  using G = device_graph_t<CSR, E>;
  vdata_t<WA, RA> data;

  __device__ __forceinline__ WA *wa_of(int v) const { return data.fetch_wa(v); }
  __device__ __forceinline__ RA *ra_of(int v) const { return data.fetch_ra(v); }

  __device__ Status filter(int v, G g) { return Inactive; }
  __device__ WA emit(int v, E *e, G g) { return *data.fetch_wa(v); }
  __device__ bool cond(int v, WA msg, G g) { return false; }
  __device__ bool comp(WA *v_data, WA msg, G g) { return true; }
  __device__ bool compAtomic(WA *v_data, WA msg, G g) { return true; }
  __device__ bool exit(int v, G g) { return false; }
};
