// Host-side graph ingest: generators and CSR builder (see host_graph.h).
#include "host_graph.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <thread>

namespace gsw_host {

int num_threads() {
  const char *env = std::getenv("GSW_HOST_THREADS");
  int t = env ? std::atoi(env) : 0;
  if (t <= 0) t = (int)std::thread::hardware_concurrency();
  if (t <= 0) t = 1;
  return std::min(t, 16);
}

template <typename F> static void run_threads(int T, F fn) {
  if (T <= 1) {
    fn(0);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(T);
  for (int t = 0; t < T; ++t) th.emplace_back(fn, t);
  for (auto &x : th) x.join();
}

// CSR build by counting sort: per-source degree counts (relaxed atomics),
// exclusive scan, scatter of (target, weight), then every row sorted and
// de-duplicated in parallel (rows split by edge count), then compacted.
static void build_one(int64_t n, int64_t ne, const int32_t *src,
                      const int32_t *dst, const uint32_t *w, bool directed,
                      bool reverse, std::vector<int64_t> &row_ptr,
                      std::vector<int32_t> &col, std::vector<uint32_t> &ww) {
  const int T = (ne < (1 << 16)) ? 1 : num_threads();
  // per-thread private counts; turned into per-thread offsets inside a row
  std::vector<std::vector<int32_t>> cnt((size_t)T);
  auto each_edge = [&](auto fn) {
    run_threads(T, [&](int t) {
      for (int64_t i = ne * t / T, e = ne * (t + 1) / T; i < e; ++i) {
        int32_t u = src[i], v = dst[i];
        if (u == v) continue;
        if (reverse) std::swap(u, v);
        fn(t, u, v, w ? w[i] : 0u);
        if (!directed) fn(t, v, u, w ? w[i] : 0u);
      }
    });
  };
  run_threads(T, [&](int t) { cnt[t].assign((size_t)n, 0); });
  each_edge([&](int t, int32_t u, int32_t, uint32_t) { ++cnt[t][u]; });
  std::vector<int64_t> raw((size_t)n + 1, 0);
  {
    // raw = exclusive scan of row totals; cnt[t][v] <- rows written by
    // threads before t
    std::vector<int64_t> tot((size_t)n, 0);
    run_threads(T, [&](int t) {
      for (int64_t v = n * t / T, e = n * (t + 1) / T; v < e; ++v) {
        int64_t run = 0;
        for (int k = 0; k < T; ++k) {
          const int32_t c = cnt[k][v];
          cnt[k][v] = (int32_t)run;
          run += c;
        }
        tot[v] = run;
      }
    });
    for (int64_t v = 0; v < n; ++v) raw[v + 1] = raw[v] + tot[v];
  }
  const int64_t M = raw[n];
  std::vector<int32_t> rc((size_t)M);
  std::vector<uint32_t> rw(w ? (size_t)M : 0);
  each_edge([&](int t, int32_t u, int32_t v, uint32_t x) {
    const int64_t p = raw[u] + cnt[t][u]++;
    rc[p] = v;
    if (w) rw[p] = x;
  });
  cnt.clear();
  cnt.shrink_to_fit();
  // rows [r0, r1) per thread, balanced by edges
  std::vector<int64_t> cut((size_t)T + 1, n);
  cut[0] = 0;
  for (int t = 1; t < T; ++t)
    cut[t] = std::lower_bound(raw.begin(), raw.end(), M * t / T) - raw.begin();
  for (int t = 1; t <= T; ++t) cut[t] = std::max(cut[t], cut[t - 1]);
  std::vector<int64_t> deg((size_t)n, 0);
  run_threads(T, [&](int t) {
    std::vector<std::pair<int32_t, uint32_t>> tmp;
    for (int64_t v = cut[t]; v < cut[t + 1]; ++v) {
      const int64_t s = raw[v], e = raw[v + 1];
      if (!w) {
        std::sort(rc.begin() + s, rc.begin() + e);
        deg[v] = std::unique(rc.begin() + s, rc.begin() + e) - (rc.begin() + s);
      } else {
        tmp.clear();
        for (int64_t j = s; j < e; ++j) tmp.push_back({rc[j], rw[j]});
        std::sort(tmp.begin(), tmp.end());
        int64_t k = 0;
        for (size_t j = 0; j < tmp.size(); ++j)
          if (j == 0 || tmp[j].first != tmp[j - 1].first) {
            rc[s + k] = tmp[j].first;
            rw[s + k] = tmp[j].second;
            ++k;
          }
        deg[v] = k;
      }
    }
  });
  row_ptr.assign((size_t)n + 1, 0);
  for (int64_t v = 0; v < n; ++v) row_ptr[v + 1] = row_ptr[v] + deg[v];
  col.resize((size_t)row_ptr[n]);
  if (w) ww.resize((size_t)row_ptr[n]);
  run_threads(T, [&](int t) {
    for (int64_t v = cut[t]; v < cut[t + 1]; ++v) {
      std::copy(rc.begin() + raw[v], rc.begin() + raw[v] + deg[v],
                col.begin() + row_ptr[v]);
      if (w)
        std::copy(rw.begin() + raw[v], rw.begin() + raw[v] + deg[v],
                  ww.begin() + row_ptr[v]);
    }
  });
}

void build_csr(int64_t n, int64_t ne, const int32_t *src, const int32_t *dst,
               const uint32_t *w, bool directed, uint64_t wkey,
               uint32_t wlim, HostGraph &g) {
  if (n < 0 || n > (int64_t)INT32_MAX) throw std::invalid_argument("n out of range");
  g = HostGraph();
  g.n = n;
  g.directed = directed;
  g.weighted = (w != nullptr) || (wkey != 0);
  for (int64_t i = 0; i < ne; ++i)
    if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n)
      throw std::invalid_argument("edge endpoint out of range");
  build_one(n, ne, src, dst, w, directed, false, g.row_ptr, g.col, g.w);
  if (directed) build_one(n, ne, src, dst, w, true, true, g.r_row_ptr, g.r_col, g.r_w);
  if (w == nullptr && wkey != 0) {
    auto fill_w = [&](const std::vector<int64_t> &rp,
                      const std::vector<int32_t> &cl, std::vector<uint32_t> &ww) {
      ww.resize(cl.size());
      const int T = num_threads();
      run_threads(T, [&](int t) {
        int64_t v0 = n * t / T, v1 = n * (t + 1) / T;
        for (int64_t v = v0; v < v1; ++v)
          for (int64_t j = rp[v]; j < rp[v + 1]; ++j) {
            uint32_t a = (uint32_t)v, c = (uint32_t)cl[j];
            ww[j] = pair_weight(wkey, std::min(a, c), std::max(a, c), wlim);
          }
      });
    };
    fill_w(g.row_ptr, g.col, g.w);
    if (directed) fill_w(g.r_row_ptr, g.r_col, g.r_w);
  }
  g.m = (int64_t)g.col.size();
}

// R-MAT (Chakrabarti et al. 2004): for each of edge_factor * 2^scale edge
// samples, `scale` quadrant choices with probabilities (a, b, c, 1-a-b-c),
// most significant bit first. Choice k of sample i reads 32 bits of draw
// (key, i*ceil(scale/2) + k/2): low half for even k, high half for odd k,
// compared against floor(p * 2^32) cumulative thresholds.
void gen_rmat(int scale, int edge_factor, double a, double b, double c,
              uint64_t seed, std::vector<int32_t> &src,
              std::vector<int32_t> &dst) {
  if (scale < 1 || scale > 30) throw std::invalid_argument("scale out of range");
  const int64_t ne = (int64_t)edge_factor << scale;
  src.resize((size_t)ne);
  dst.resize((size_t)ne);
  const uint64_t key = stream_key(seed, STREAM_RMAT);
  const uint64_t t1 = (uint64_t)std::floor(a * 4294967296.0);
  const uint64_t t2 = (uint64_t)std::floor((a + b) * 4294967296.0);
  const uint64_t t3 = (uint64_t)std::floor((a + b + c) * 4294967296.0);
  const int per = (scale + 1) / 2;
  const int T = num_threads();
  run_threads(T, [&](int t) {
    for (int64_t i = ne * t / T, e = ne * (t + 1) / T; i < e; ++i) {
      uint32_t u = 0, v = 0;
      uint64_t r = 0;
      for (int k = 0; k < scale; ++k) {
        if ((k & 1) == 0) r = draw(key, (uint64_t)i * per + (k >> 1));
        uint64_t x = (k & 1) ? (r >> 32) : (r & 0xffffffffull);
        uint32_t ubit = 0, vbit = 0;
        if (x < t1) {
        } else if (x < t2) {
          vbit = 1;
        } else if (x < t3) {
          ubit = 1;
        } else {
          ubit = 1;
          vbit = 1;
        }
        u = (u << 1) | ubit;
        v = (v << 1) | vbit;
      }
      src[i] = (int32_t)u;
      dst[i] = (int32_t)v;
    }
  });
}

// rows x cols 4-neighbour grid, id = r*cols + c. Horizontal edge (r,c)-(r,c+1)
// has id r*(cols-1)+c; vertical edge (r,c)-(r+1,c) has id
// rows*(cols-1) + r*cols + c. Edge id e is deleted when
// draw(delete_key, e) < p_delete * 2^64. Then round(long_range_frac * kept)
// extra edges with endpoints below(draw(long_key, 2j), n),
// below(draw(long_key, 2j+1), n).
void gen_grid2d(int64_t rows, int64_t cols, double p_delete,
                double long_range_frac, uint64_t seed,
                std::vector<int32_t> &src, std::vector<int32_t> &dst) {
  const uint64_t dkey = stream_key(seed, STREAM_GRID_DELETE);
  const uint64_t lkey = stream_key(seed, STREAM_GRID_LONG);
  const double lim = p_delete * 18446744073709551616.0;
  const uint64_t thr = lim >= 18446744073709551615.0 ? ~0ull : (uint64_t)lim;
  src.clear();
  dst.clear();
  const int64_t nh = rows * (cols - 1);
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t c = 0; c + 1 < cols; ++c) {
      int64_t e = r * (cols - 1) + c;
      if (draw(dkey, (uint64_t)e) < thr) continue;
      src.push_back((int32_t)(r * cols + c));
      dst.push_back((int32_t)(r * cols + c + 1));
    }
  for (int64_t r = 0; r + 1 < rows; ++r)
    for (int64_t c = 0; c < cols; ++c) {
      int64_t e = nh + r * cols + c;
      if (draw(dkey, (uint64_t)e) < thr) continue;
      src.push_back((int32_t)(r * cols + c));
      dst.push_back((int32_t)((r + 1) * cols + c));
    }
  const int64_t n = rows * cols;
  const int64_t kept = (int64_t)src.size();
  const int64_t L = (int64_t)std::llround(long_range_frac * (double)kept);
  for (int64_t j = 0; j < L; ++j) {
    src.push_back((int32_t)below(draw(lkey, 2 * (uint64_t)j), (uint64_t)n));
    dst.push_back((int32_t)below(draw(lkey, 2 * (uint64_t)j + 1), (uint64_t)n));
  }
}

// Erdos-Renyi G(n, M) with M = round(n * mean_degree / 2) endpoint pairs
// drawn uniformly (self-loops and repeats are removed by the CSR builder).
void gen_er(int64_t n, double mean_degree, uint64_t seed,
            std::vector<int32_t> &src, std::vector<int32_t> &dst) {
  const uint64_t key = stream_key(seed, STREAM_ER);
  const int64_t M = (int64_t)std::llround((double)n * mean_degree / 2.0);
  src.resize((size_t)M);
  dst.resize((size_t)M);
  const int T = (M < (1 << 16)) ? 1 : num_threads();
  run_threads(T, [&](int t) {
    for (int64_t j = M * t / T, e = M * (t + 1) / T; j < e; ++j) {
      src[j] = (int32_t)below(draw(key, 2 * (uint64_t)j), (uint64_t)n);
      dst[j] = (int32_t)below(draw(key, 2 * (uint64_t)j + 1), (uint64_t)n);
    }
  });
}

// Barabasi-Albert: a clique on m+1 vertices, then every new vertex t attaches
// to m distinct earlier vertices drawn with probability proportional to
// degree (uniform pick from the endpoint list of all edges so far; repeats
// rejected). Draws come from draw(key, ctr) with one counter per attempt.
void gen_ba(int64_t n, int m, uint64_t seed, std::vector<int32_t> &src,
            std::vector<int32_t> &dst) {
  if (m < 1) throw std::invalid_argument("m must be >= 1");
  const uint64_t key = stream_key(seed, STREAM_BA);
  src.clear();
  dst.clear();
  std::vector<int32_t> ends;
  const int64_t m0 = std::min<int64_t>(n, (int64_t)m + 1);
  for (int64_t a = 0; a < m0; ++a)
    for (int64_t b = a + 1; b < m0; ++b) {
      src.push_back((int32_t)a);
      dst.push_back((int32_t)b);
      ends.push_back((int32_t)a);
      ends.push_back((int32_t)b);
    }
  uint64_t ctr = 0;
  std::vector<int32_t> chosen;
  for (int64_t t = m0; t < n; ++t) {
    chosen.clear();
    while ((int)chosen.size() < m) {
      int32_t x = ends[below(draw(key, ctr++), ends.size())];
      if (std::find(chosen.begin(), chosen.end(), x) == chosen.end())
        chosen.push_back(x);
    }
    for (int32_t x : chosen) {
      src.push_back((int32_t)t);
      dst.push_back(x);
      ends.push_back((int32_t)t);
      ends.push_back(x);
    }
  }
}

}  // namespace gsw_host

namespace gsw_host {

// Instance i: u1, u2 = the top 53 bits of draw(batch_key, 2i) and
// draw(batch_key, 2i + 1) as uniforms in [0, 1); n_i = floor(exp(ln nmin +
// u1 (ln nmax - ln nmin))), d_i = dmin + u2 (dmax - dmin); its edges are
// gen_er(n_i, d_i, seed_i) with seed_i = stream_key(seed, 1000 + i), shifted
// by the instance's vertex offset. Instances first .. first+count-1 are
// built, numbered from vertex 0: a batch built in consecutive ranges is the
// whole batch cut into pieces.
void gen_er_batch(int64_t first, int count, int64_t nmin, int64_t nmax,
                  double dmin, double dmax, uint64_t seed,
                  std::vector<int32_t> &src, std::vector<int32_t> &dst,
                  std::vector<int64_t> &offsets) {
  if (first < 0 || count < 1 || nmin < 1 || nmax < nmin)
    throw std::invalid_argument("invalid batch parameters");
  const uint64_t key = stream_key(seed, STREAM_BATCH);
  std::vector<int64_t> ns(count);
  std::vector<double> ds(count);
  offsets.assign(count + 1, 0);
  const double l0 = std::log((double)nmin), l1 = std::log((double)nmax);
  for (int i = 0; i < count; ++i) {
    const uint64_t k = (uint64_t)(first + i);
    const double u1 = (double)(draw(key, 2 * k) >> 11) * 0x1.0p-53;
    const double u2 = (double)(draw(key, 2 * k + 1) >> 11) * 0x1.0p-53;
    ns[i] = std::max<int64_t>(1, (int64_t)std::floor(std::exp(l0 + u1 * (l1 - l0))));
    ns[i] = std::min(ns[i], nmax);
    ds[i] = dmin + u2 * (dmax - dmin);
    offsets[i + 1] = offsets[i] + ns[i];
  }
  if (offsets[count] > (int64_t)INT32_MAX)
    throw std::invalid_argument("batch exceeds int32 vertex ids");
  // edge counts per instance, then every instance generated in parallel
  std::vector<int64_t> eo(count + 1, 0);
  for (int i = 0; i < count; ++i)
    eo[i + 1] = eo[i] + (int64_t)std::llround((double)ns[i] * ds[i] / 2.0);
  src.resize((size_t)eo[count]);
  dst.resize((size_t)eo[count]);
  const int T = num_threads();
  run_threads(T, [&](int t) {
    std::vector<int32_t> s, d;
    for (int i = t; i < count; i += T) {
      gen_er(ns[i], ds[i], stream_key(seed, 1000 + (uint64_t)(first + i)), s, d);
      for (size_t j = 0; j < s.size(); ++j) {
        src[eo[i] + j] = (int32_t)(s[j] + offsets[i]);
        dst[eo[i] + j] = (int32_t)(d[j] + offsets[i]);
      }
    }
  });
}

}  // namespace gsw_host
