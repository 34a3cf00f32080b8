// Host-side graph ingest: seeded synthetic generators and the CSR builder.
//
// CSR semantics follow the reference's host_graph_t<CSR>::build + rinse
// (src/utils/fileIO.cuh:210-332, 399-443): self-loops dropped
// (fileIO.cuh:128-129), undirected input stored in both directions
// (fileIO.cuh:227-228, 291-296), duplicates removed and every adjacency list
// sorted ascending (rinse, fileIO.cuh:407-417). Unlike rinse, isolated
// vertices keep their ids (no renumbering), so generated vertex ids are
// stable. Offsets are int64 (row_ptr[n+1]) rather than int32 start_pos[n].
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gsw_host {

// Counter-based RNG: SplitMix64's output mix applied to key + (ctr+1)*phi.
// Every random draw is a pure function of (key, counter), so generation is
// order-independent, parallel, and reproducible from Python.
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return mix64(mix64(seed) + stream * 0xD1B54A32D192ED03ull);
}
inline uint64_t draw(uint64_t key, uint64_t ctr) {
  return mix64(key + (ctr + 1) * 0x9E3779B97F4A7C15ull);
}
// uniform integer in [0, n) from a 64-bit draw (multiply-high)
inline uint64_t below(uint64_t r, uint64_t n) {
  return (uint64_t)(((unsigned __int128)r * n) >> 64);
}

enum : uint64_t {
  STREAM_RMAT = 1,
  STREAM_WEIGHT = 2,
  STREAM_GRID_DELETE = 3,
  STREAM_GRID_LONG = 4,
  STREAM_ER = 5,
  STREAM_BA = 6,
  STREAM_BATCH = 7,
};

// Symmetric edge weight in [1, lim] for the canonical pair (lo, hi).
inline uint32_t pair_weight(uint64_t wkey, uint32_t lo, uint32_t hi,
                            uint32_t lim) {
  uint64_t r = draw(wkey, ((uint64_t)lo << 32) | hi);
  return 1u + (uint32_t)(((r >> 32) * (uint64_t)lim) >> 32);
}

struct HostGraph {
  int64_t n = 0;
  int64_t m = 0;  // directed CSR entries
  bool directed = false;
  bool weighted = false;
  std::vector<int64_t> row_ptr;  // n+1
  std::vector<int32_t> col;      // m
  std::vector<uint32_t> w;       // m or empty
  // reverse CSR (directed graphs only)
  std::vector<int64_t> r_row_ptr;
  std::vector<int32_t> r_col;
  std::vector<uint32_t> r_w;
};

// Build CSR from an edge list. Undirected: both directions stored.
// Weighted input with duplicates keeps the minimum weight of the duplicates.
// If w == nullptr and wkey != 0, weights are drawn with pair_weight().
void build_csr(int64_t n, int64_t ne, const int32_t *src, const int32_t *dst,
               const uint32_t *w, bool directed, uint64_t wkey,
               uint32_t wlim, HostGraph &g);

void gen_rmat(int scale, int edge_factor, double a, double b, double c,
              uint64_t seed, std::vector<int32_t> &src,
              std::vector<int32_t> &dst);
void gen_grid2d(int64_t rows, int64_t cols, double p_delete,
                double long_range_frac, uint64_t seed,
                std::vector<int32_t> &src, std::vector<int32_t> &dst);
void gen_er(int64_t n, double mean_degree, uint64_t seed,
            std::vector<int32_t> &src, std::vector<int32_t> &dst);
void gen_ba(int64_t n, int m, uint64_t seed, std::vector<int32_t> &src,
            std::vector<int32_t> &dst);
// `count` Erdos-Renyi instances as one disjoint union (see gsw_gen_er_batch);
// offsets receives count + 1 vertex offsets.
void gen_er_batch(int64_t first, int count, int64_t nmin, int64_t nmax,
                  double dmin, double dmax, uint64_t seed,
                  std::vector<int32_t> &src, std::vector<int32_t> &dst,
                  std::vector<int64_t> &offsets);

// Matrix Market / edge-list file -> CSR (see mtx.cpp for the format rules).
void read_mtx(const std::string &path, bool directed, bool with_weight,
              bool with_header, bool rinse, HostGraph &g);

// Binary CSR cache: save / load a HostGraph (see mtx.cpp for the layout).
void save_csr(const HostGraph &g, const std::string &path);
void load_csr(const std::string &path, HostGraph &g);

int num_threads();

}  // namespace gsw_host
