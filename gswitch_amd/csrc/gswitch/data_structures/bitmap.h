// Frontier bitmaps with 64-bit words (one word per wavefront ballot).
//
// Semantics follow bitmap_t (src/data_structures/bitmap.cuh:116-164):
//   active   : bit = 1 for Active vertices
//   inactive : bit = 0 for Inactive vertices, 1 otherwise
//   visited  : de-duplication marks for the expand phase
// The reference stores byte words and marks `visited` with an unguarded
// read-modify-write ("no consistency guarantee", bitmap.cuh:71-99); here every
// mark is a 64-bit atomicOr, so de-duplication is exact.
#pragma once
#include "../utils/common.h"

struct bits_t {
  uint64_t *dg_bits = nullptr;
  int64_t n_words = 0;

  int64_t build(int64_t nv) {
    n_words = (nv + 63) >> 6;
    H_ERR(hipMalloc((void **)&dg_bits, sizeof(uint64_t) * (n_words + 1)));
    H_ERR(hipMemset(dg_bits, 0, sizeof(uint64_t) * (n_words + 1)));
    return sizeof(uint64_t) * n_words;
  }
  void release() {
    hipFree(dg_bits);
    dg_bits = nullptr;
  }

  __device__ __forceinline__ uint64_t word(int64_t w) const { return dg_bits[w]; }
  __device__ __forceinline__ bool test(int v) const {
    return (dg_bits[v >> 6] >> (v & 63)) & 1ull;
  }
  // set bit v; true if this call flipped it from 0 to 1
  __device__ __forceinline__ bool test_and_set(int v) {
    const uint64_t bit = 1ull << (v & 63);
    if (dg_bits[v >> 6] & bit) return false;
    return !(atomicOr((unsigned long long *)&dg_bits[v >> 6],
                      (unsigned long long)bit) &
             bit);
  }
  __device__ __forceinline__ void set_word(int64_t w, uint64_t x) { dg_bits[w] = x; }
};

struct bitmap_t {
  bits_t visited, active, inactive;
  int64_t nv = 0;

  int64_t build(int64_t n) {
    nv = n;
    return visited.build(n) + active.build(n) + inactive.build(n);
  }
  void release() {
    visited.release();
    active.release();
    inactive.release();
  }

  __device__ __forceinline__ bool is_active(int v) const { return active.test(v); }
  __device__ __forceinline__ bool is_inactive(int v) const { return !inactive.test(v); }
  __device__ __forceinline__ bool is_valid(int v) const {
    return is_active(v) || is_inactive(v);
  }
  // true if v was not marked yet (and marks it)
  __device__ __forceinline__ bool mark_duplicate_lite(int v) {
    return visited.test_and_set(v);
  }
  __device__ __forceinline__ bool mark_duplicate_atomic(int v) {
    return visited.test_and_set(v);
  }
};
