// Graph ingest from Matrix Market / whitespace edge-list files.
//
// Parsing follows edgelist_t::read_mtx (src/utils/fileIO.cuh:73-146):
//   * lines are skipped until one starts with a digit (comments, the
//     %%MatrixMarket banner, blank lines);
//   * with_header: that first numeric line ("rows cols nnz") is ignored;
//     without it, it is the first edge;
//   * each record is "v0 v1 [w]": two integer ids, then the weight when
//     with_weight, otherwise the rest of the line is ignored;
//   * a record whose line is not ended by a newline (the last line of a
//     file without a trailing newline) is dropped, as the reference's
//     stream loop drops it at end-of-file; its ids still count for the id
//     range unless the file ends right after v1;
//   * self-loops are dropped after the ids count for the range;
//   * ids are shifted by the smallest id seen (vmin), n = vmax - vmin + 1;
//   * no edge or no vertex is an error (the reference exits).
// The CSR is then built by build_csr (rows sorted, duplicates removed,
// undirected input stored both ways). With rinse, vertices of degree zero
// are removed and the others renumbered in id order, as the reference's
// rinse (fileIO.cuh:399-443) does for unweighted graphs.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "host_graph.h"

namespace gsw_host {

namespace {

struct cursor_t {
  const char *p, *end;
  bool eof() const { return p >= end; }
  void skip_space() {
    while (p < end && std::isspace((unsigned char)*p)) ++p;
  }
  // next integer (whitespace, newlines included, skipped first); false at
  // end of input or on a token that does not start an integer
  bool next_int(int64_t &x) {
    skip_space();
    if (p >= end) return false;
    bool neg = false;
    if (*p == '-' || *p == '+') {
      neg = (*p == '-');
      ++p;
    }
    if (p >= end || !std::isdigit((unsigned char)*p)) return false;
    int64_t v = 0;
    while (p < end && std::isdigit((unsigned char)*p)) {
      v = v * 10 + (*p - '0');
      if (v > (int64_t)INT32_MAX + 1) throw std::invalid_argument("vertex id or weight out of range");
      ++p;
    }
    x = neg ? -v : v;
    return true;
  }
  // skip to just past the next newline; false if the input ends first
  bool skip_line() {
    while (p < end && *p != '\n') ++p;
    if (p >= end) return false;
    ++p;
    return true;
  }
};

}  // namespace

void read_mtx(const std::string &path, bool directed, bool with_weight,
              bool with_header, bool rinse, HostGraph &g) {
  FILE *fp = std::fopen(path.c_str(), "rb");
  if (!fp) throw std::invalid_argument("can not open file " + path);
  std::string buf;
  {
    char tmp[1 << 16];
    size_t k;
    while ((k = std::fread(tmp, 1, sizeof(tmp), fp)) > 0) buf.append(tmp, k);
    std::fclose(fp);
  }
  cursor_t c{buf.data(), buf.data() + buf.size()};
  // skip lines until one starts with a digit
  while (!c.eof() && !std::isdigit((unsigned char)*c.p))
    if (!c.skip_line()) break;
  if (with_header && !c.eof()) c.skip_line();

  std::vector<int32_t> src, dst;
  std::vector<uint32_t> w;
  int64_t vmin = INT32_MAX, vmax = -1;
  bool any = false;
  while (true) {
    int64_t v0, v1;
    if (!c.next_int(v0) || !c.next_int(v1)) break;
    if (c.eof()) break;  // "v0 v1" ran to the end of the file: not counted
    if (v0 < 0 || v1 < 0) throw std::invalid_argument("negative vertex id");
    vmin = std::min(vmin, std::min(v0, v1));
    vmax = std::max(vmax, std::max(v0, v1));
    any = true;
    int64_t wt = 0;
    if (with_weight) {
      if (!c.next_int(wt)) break;
      if (c.eof()) break;  // weight ran to the end of the file
      if (*c.p == '.' || *c.p == 'e' || *c.p == 'E')
        throw std::invalid_argument("edge weights must be integers");
      if (wt < 0) throw std::invalid_argument("negative edge weight");
    } else if (!c.skip_line()) {
      break;
    }
    if (v0 == v1) continue;
    src.push_back((int32_t)v0);
    dst.push_back((int32_t)v1);
    if (with_weight) w.push_back((uint32_t)wt);
  }
  if (!any || src.empty()) throw std::invalid_argument("graph has no edges");
  const int64_t n = vmax - vmin + 1;
  if (n > (int64_t)INT32_MAX) throw std::invalid_argument("id range exceeds int32");
  for (size_t i = 0; i < src.size(); ++i) {
    src[i] = (int32_t)(src[i] - vmin);
    dst[i] = (int32_t)(dst[i] - vmin);
  }
  build_csr(n, (int64_t)src.size(), src.data(), dst.data(),
            with_weight ? w.data() : nullptr, directed, 0, 64, g);
  if (!rinse) return;
  if (directed) throw std::invalid_argument("rinse applies to undirected graphs only");
  // drop degree-zero vertices, renumber the rest in id order
  std::vector<int32_t> rename((size_t)g.n, -1);
  int64_t top = 0;
  for (int64_t v = 0; v < g.n; ++v)
    if (g.row_ptr[v + 1] > g.row_ptr[v]) rename[v] = (int32_t)top++;
  if (top == g.n) return;
  std::vector<int64_t> rp((size_t)top + 1, 0);
  for (int64_t v = 0; v < g.n; ++v)
    if (rename[v] >= 0) rp[rename[v] + 1] = g.row_ptr[v + 1] - g.row_ptr[v];
  for (int64_t v = 0; v < top; ++v) rp[v + 1] += rp[v];
  for (auto &u : g.col) u = rename[u];  // rows keep their order
  g.row_ptr.swap(rp);
  g.n = top;
}

}  // namespace gsw_host

namespace gsw_host {

// Binary CSR cache (the reference's store_to_binfile / load_from_binfile,
// src/utils/fileIO.cuh:334-387, which dump raw int arrays): here an 8-byte
// tag "GSWCSR01", then int64 n, int64 m, int32 directed, int32 weighted,
// then row_ptr int64[n+1], col int32[m], w uint32[m] (weighted), and for a
// directed graph the reverse row_ptr / col / w the same way.
namespace {
const char kTag[8] = {'G', 'S', 'W', 'C', 'S', 'R', '0', '1'};

struct file_t {
  FILE *f;
  explicit file_t(FILE *x) : f(x) {}
  ~file_t() {
    if (f) std::fclose(f);
  }
};
template <typename T> void put(FILE *f, const T *p, size_t k) {
  if (k && std::fwrite(p, sizeof(T), k, f) != k)
    throw std::runtime_error("write failed");
}
template <typename T> void get(FILE *f, T *p, size_t k) {
  if (k && std::fread(p, sizeof(T), k, f) != k)
    throw std::invalid_argument("truncated CSR cache file");
}
}  // namespace

void save_csr(const HostGraph &g, const std::string &path) {
  file_t fp(std::fopen(path.c_str(), "wb"));
  if (!fp.f) throw std::invalid_argument("can not create file " + path);
  const int64_t head[2] = {g.n, g.m};
  const int32_t flags[2] = {g.directed ? 1 : 0, g.weighted ? 1 : 0};
  put(fp.f, kTag, 8);
  put(fp.f, head, 2);
  put(fp.f, flags, 2);
  put(fp.f, g.row_ptr.data(), (size_t)g.n + 1);
  put(fp.f, g.col.data(), (size_t)g.m);
  if (g.weighted) put(fp.f, g.w.data(), (size_t)g.m);
  if (g.directed) {
    put(fp.f, g.r_row_ptr.data(), (size_t)g.n + 1);
    put(fp.f, g.r_col.data(), (size_t)g.m);
    if (g.weighted) put(fp.f, g.r_w.data(), (size_t)g.m);
  }
}

void load_csr(const std::string &path, HostGraph &g) {
  file_t fp(std::fopen(path.c_str(), "rb"));
  if (!fp.f) throw std::invalid_argument("can not open file " + path);
  char tag[8];
  get(fp.f, tag, 8);
  if (std::string(tag, 8) != std::string(kTag, 8))
    throw std::invalid_argument("not a CSR cache file: " + path);
  int64_t head[2];
  int32_t flags[2];
  get(fp.f, head, 2);
  get(fp.f, flags, 2);
  if (head[0] < 0 || head[0] > (int64_t)INT32_MAX || head[1] < 0)
    throw std::invalid_argument("corrupt CSR cache header");
  g = HostGraph();
  g.n = head[0];
  g.m = head[1];
  g.directed = flags[0] != 0;
  g.weighted = flags[1] != 0;
  {
    // the header's sizes must match the file before anything is allocated
    const long here = std::ftell(fp.f);
    if (here < 0 || std::fseek(fp.f, 0, SEEK_END) != 0)
      throw std::invalid_argument("can not size CSR cache file " + path);
    const long end = std::ftell(fp.f);
    if (end < 0 || std::fseek(fp.f, here, SEEK_SET) != 0)
      throw std::invalid_argument("can not size CSR cache file " + path);
    const double side = 8.0 * ((double)g.n + 1) +
                        (g.weighted ? 8.0 : 4.0) * (double)g.m;
    const double want = (double)here + side * (g.directed ? 2 : 1);
    if ((double)end != want)
      throw std::invalid_argument("CSR cache file size does not match its header");
  }
  auto read_side = [&](std::vector<int64_t> &rp, std::vector<int32_t> &cl,
                       std::vector<uint32_t> &w) {
    rp.resize((size_t)g.n + 1);
    cl.resize((size_t)g.m);
    get(fp.f, rp.data(), rp.size());
    get(fp.f, cl.data(), cl.size());
    if (g.weighted) {
      w.resize((size_t)g.m);
      get(fp.f, w.data(), w.size());
    }
    if (rp[0] != 0 || rp[g.n] != g.m)
      throw std::invalid_argument("corrupt CSR cache offsets");
    for (int64_t v = 0; v < g.n; ++v)
      if (rp[v + 1] < rp[v]) throw std::invalid_argument("corrupt CSR cache offsets");
    for (int32_t u : cl)
      if (u < 0 || u >= g.n) throw std::invalid_argument("corrupt CSR cache ids");
    // the writer stores sorted rows; kernels that split rows by id range
    // (the TWOD slices) rely on it
    for (int64_t v = 0; v < g.n; ++v)
      if (!std::is_sorted(cl.begin() + rp[v], cl.begin() + rp[v + 1]))
        throw std::invalid_argument("CSR cache rows are not sorted");
  };
  read_side(g.row_ptr, g.col, g.w);
  if (g.directed) read_side(g.r_row_ptr, g.r_col, g.r_w);
}

}  // namespace gsw_host
