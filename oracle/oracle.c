/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Single-threaded, as the reference's validation routines are.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ---- BFS: application/bfs.cu:38-62 --------------------------------------
 * FIFO queue from the root; a neighbour gets lvl[v]+1 the first time it is
 * seen. */
void orc_bfs(int64_t n, const int64_t *row_ptr, const int32_t *col,
             int32_t root, int32_t *lvl) {
  for (int64_t i = 0; i < n; ++i) lvl[i] = -1;
  if (n <= 0 || root < 0 || root >= n) return;
  int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  int64_t head = 0, tail = 0;
  lvl[root] = 0;
  q[tail++] = root;
  while (head < tail) {
    int32_t v = q[head++];
    for (int64_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j) {
      int32_t u = col[j];
      if (lvl[u] == -1) {
        lvl[u] = lvl[v] + 1;
        q[tail++] = u;
      }
    }
  }
  free(q);
}

/* ---- SPFA: application/sssp.cu:55-89 -------------------------------------
 * Ring-buffer queue; in_q[] keeps a vertex at most once in the queue, so a
 * capacity of n suffices. */
/* Path lengths saturate at UINT32_MAX (the unreached sentinel): a path
 * longer than 2^32 - 1 is treated as no path, where a wrapping add would
 * hand out a short distance. */
static uint32_t sat_add(uint32_t a, uint32_t b) {
  uint32_t s = a + b;
  return s < a ? UINT32_MAX : s;
}

void orc_spfa(int64_t n, const int64_t *row_ptr, const int32_t *col,
              const uint32_t *w, int32_t root, uint32_t *dist) {
  for (int64_t i = 0; i < n; ++i) dist[i] = UINT32_MAX;
  if (n <= 0 || root < 0 || root >= n) return;
  int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  char *in_q = (char *)calloc((size_t)n, 1);
  int64_t head = 0, count = 0;
  dist[root] = 0;
  q[0] = root;
  count = 1;
  in_q[root] = 1;
  while (count > 0) {
    int32_t v = q[head];
    head = (head + 1 == n) ? 0 : head + 1;
    --count;
    in_q[v] = 0;
    uint32_t dv = dist[v];
    for (int64_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j) {
      int32_t u = col[j];
      uint32_t nd = sat_add(dv, w[j]);
      if (dist[u] > nd) {
        dist[u] = nd;
        if (!in_q[u]) {
          int64_t t = head + count;
          if (t >= n) t -= n;
          q[t] = u;
          ++count;
          in_q[u] = 1;
        }
      }
    }
  }
  free(q);
  free(in_q);
}

/* ---- Dijkstra (textbook, binary min-heap keyed by distance) ------------- */
typedef struct {
  uint32_t d;
  int32_t v;
} heap_item;

static void heap_push(heap_item *h, int64_t *sz, heap_item it) {
  int64_t i = (*sz)++;
  h[i] = it;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (h[p].d <= h[i].d) break;
    heap_item t = h[p];
    h[p] = h[i];
    h[i] = t;
    i = p;
  }
}

static heap_item heap_pop(heap_item *h, int64_t *sz) {
  heap_item top = h[0];
  h[0] = h[--(*sz)];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = l + 1, s = i;
    if (l < *sz && h[l].d < h[s].d) s = l;
    if (r < *sz && h[r].d < h[s].d) s = r;
    if (s == i) break;
    heap_item t = h[s];
    h[s] = h[i];
    h[i] = t;
    i = s;
  }
  return top;
}

void orc_dijkstra(int64_t n, const int64_t *row_ptr, const int32_t *col,
                  const uint32_t *w, int32_t root, uint32_t *dist) {
  for (int64_t i = 0; i < n; ++i) dist[i] = UINT32_MAX;
  if (n <= 0 || root < 0 || root >= n) return;
  int64_t m = row_ptr[n];
  heap_item *h = (heap_item *)malloc(sizeof(heap_item) * (size_t)(m + 1));
  char *done = (char *)calloc((size_t)n, 1);
  int64_t sz = 0;
  dist[root] = 0;
  heap_item it = {0, root};
  heap_push(h, &sz, it);
  while (sz > 0) {
    heap_item top = heap_pop(h, &sz);
    int32_t v = top.v;
    if (done[v]) continue;
    done[v] = 1;
    for (int64_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j) {
      int32_t u = col[j];
      uint32_t nd = sat_add(top.d, w[j]);
      if (nd < dist[u]) {
        dist[u] = nd;
        heap_item ni = {nd, u};
        heap_push(h, &sz, ni);
      }
    }
  }
  free(h);
  free(done);
}

void orc_bellman_ford(int64_t n, const int64_t *row_ptr, const int32_t *col,
                      const uint32_t *w, int32_t root, uint32_t *dist) {
  for (int64_t i = 0; i < n; ++i) dist[i] = UINT32_MAX;
  if (n <= 0 || root < 0 || root >= n) return;
  dist[root] = 0;
  for (int64_t round = 0; round + 1 < n; ++round) {
    int changed = 0;
    for (int64_t v = 0; v < n; ++v) {
      if (dist[v] == UINT32_MAX) continue;
      for (int64_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j) {
        uint32_t nd = sat_add(dist[v], w[j]);
        if (nd < dist[col[j]]) {
          dist[col[j]] = nd;
          changed = 1;
        }
      }
    }
    if (!changed) break;
  }
}

/* ---- CC: application/cc.cu:34-66 -----------------------------------------
 * Ascending ids; an id still labelled with itself starts a BFS that labels
 * every vertex still carrying its own id with the starting id. */
void orc_cc(int64_t n, const int64_t *row_ptr, const int32_t *col,
            int32_t *label) {
  for (int64_t i = 0; i < n; ++i) label[i] = (int32_t)i;
  if (n <= 0) return;
  int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    if (label[i] != i) continue;
    int32_t c = (int32_t)i;
    int64_t head = 0, tail = 0;
    q[tail++] = (int32_t)i;
    while (head < tail) {
      int32_t v = q[head++];
      for (int64_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j) {
        int32_t u = col[j];
        if (label[u] == u && u != c) {
          label[u] = c;
          q[tail++] = u;
        }
      }
    }
  }
  free(q);
}

static int32_t uf_find(int32_t *p, int32_t x) {
  while (p[x] != x) {
    p[x] = p[p[x]];
    x = p[x];
  }
  return x;
}

void orc_cc_unionfind(int64_t n, const int64_t *row_ptr, const int32_t *col,
                      int32_t *label) {
  for (int64_t i = 0; i < n; ++i) label[i] = (int32_t)i;
  for (int64_t v = 0; v < n; ++v) {
    for (int64_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j) {
      int32_t a = uf_find(label, (int32_t)v), b = uf_find(label, col[j]);
      if (a < b)
        label[b] = a;
      else if (b < a)
        label[a] = b;
    }
  }
  for (int64_t v = 0; v < n; ++v) label[v] = uf_find(label, (int32_t)v);
}

/* ---- PageRank FP64 power iteration (form of application/pagerank.cu:19-32) */
void orc_pagerank(int64_t n, const int64_t *row_ptr, const int32_t *col,
                  double damping, int iters, double *pr) {
  double *contrib = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int64_t v = 0; v < n; ++v) pr[v] = 1.0 - damping;
  for (int it = 0; it < iters; ++it) {
    for (int64_t v = 0; v < n; ++v) {
      int64_t od = row_ptr[v + 1] - row_ptr[v];
      contrib[v] = od ? pr[v] / (double)od : 0.0;
    }
    for (int64_t v = 0; v < n; ++v) {
      double s = 0.0;
      for (int64_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j) s += contrib[col[j]];
      pr[v] = (1.0 - damping) + damping * s;
    }
  }
  free(contrib);
}

/* ---- graph colouring: application/gc.cu:15-98 -----------------------------
 * Jones-Plassmann rounds r = 1, 2, ...: from the colours at the start of the
 * round, every uncoloured v finds the largest id among its uncoloured
 * neighbours; v takes colour r when its own id is larger. */
void orc_gc(int64_t n, const int64_t *row_ptr, const int32_t *col,
            int32_t *color) {
  int32_t *maxid = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t v = 0; v < n; ++v) color[v] = -1;
  int64_t left = n;
  for (int32_t r = 1; left > 0; ++r) {
    for (int64_t v = 0; v < n; ++v) {
      if (color[v] != -1) continue;
      int32_t m = -1;
      for (int64_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j) {
        const int32_t u = col[j];
        if (color[u] == -1 && u > m) m = u;
      }
      maxid[v] = m;
    }
    for (int64_t v = 0; v < n; ++v)
      if (color[v] == -1 && maxid[v] < v) {
        color[v] = r;
        --left;
      }
  }
  free(maxid);
}

/* ---- Brandes BC from one root: application/bc.cu:162-214 ----------------
 * Forward: BFS; every edge v->u with level[u] == level[v]+1 adds nsp[v] to
 * nsp[u]. Backward, in reverse BFS order: dep[v] *= nsp[v]; every edge v->u
 * with level[u]+1 == level[v] adds (1+dep[v])/nsp[v] to dep[u].
 * Before v is popped, dep[v] holds sum_w (1+dep[w])/nsp[w] over its BFS
 * successors w; the multiplication turns it into Brandes' dependency
 * delta(v) = sum_w nsp[v]/nsp[w] * (1 + delta(w)). */
void orc_bc(int64_t n, const int64_t *row_ptr, const int32_t *col,
            int32_t root, int32_t *level, double *nsp, double *dep) {
  for (int64_t i = 0; i < n; ++i) {
    level[i] = -1;
    nsp[i] = 0.0;
    dep[i] = 0.0;
  }
  if (n <= 0 || root < 0 || root >= n) return;
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  int64_t head = 0, tail = 0;
  nsp[root] = 1.0;
  level[root] = 0;
  order[tail++] = root;
  while (head < tail) {
    int32_t v = order[head++];
    for (int64_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j) {
      int32_t u = col[j];
      if (level[u] == -1) {
        level[u] = level[v] + 1;
        order[tail++] = u;
      }
      if (level[u] == level[v] + 1) nsp[u] += nsp[v];
    }
  }
  for (int64_t k = tail - 1; k >= 0; --k) {
    int32_t v = order[k];
    dep[v] *= nsp[v];
    for (int64_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j) {
      int32_t u = col[j];
      if (level[u] + 1 == level[v]) dep[u] += (1.0 + dep[v]) / nsp[v];
    }
  }
  free(order);
}

/* ---- 64-source reachability: nsrc independent BFS runs ------------------ */
void orc_msbfs64(int64_t n, const int64_t *row_ptr, const int32_t *col,
                 const int32_t *sources, int nsrc, uint64_t *mask) {
  int32_t *lvl = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  memset(mask, 0, sizeof(uint64_t) * (size_t)n);
  for (int s = 0; s < nsrc && s < 64; ++s) {
    orc_bfs(n, row_ptr, col, sources[s], lvl);
    for (int64_t v = 0; v < n; ++v)
      if (lvl[v] >= 0) mask[v] |= (1ull << s);
  }
  free(lvl);
}

int64_t orc_traversed_edges(int64_t n, const int64_t *row_ptr,
                            const int32_t *reached_level) {
  int64_t s = 0;
  for (int64_t v = 0; v < n; ++v)
    if (reached_level[v] >= 0) s += row_ptr[v + 1] - row_ptr[v];
  return s / 2;
}

/* ---- learned-selector network (FP64) ------------------------------------ */
enum { MLP_IN = 12, MLP_H = 64, MLP_OUT = 30 };
enum {
  MLP_W1 = 0,
  MLP_B1 = MLP_W1 + MLP_IN * MLP_H,
  MLP_W2 = MLP_B1 + MLP_H,
  MLP_B2 = MLP_W2 + MLP_H * MLP_H,
  MLP_W3 = MLP_B2 + MLP_H,
  MLP_B3 = MLP_W3 + MLP_H * MLP_OUT,
  MLP_NP = MLP_B3 + MLP_OUT
};

static void mlp_one(const double *w, const double *x, double *h1, double *h2,
                    double *y) {
  for (int j = 0; j < MLP_H; ++j) {
    double a = w[MLP_B1 + j];
    for (int k = 0; k < MLP_IN; ++k) a += x[k] * w[MLP_W1 + k * MLP_H + j];
    h1[j] = a > 0 ? a : 0;
  }
  for (int j = 0; j < MLP_H; ++j) {
    double a = w[MLP_B2 + j];
    for (int k = 0; k < MLP_H; ++k) a += h1[k] * w[MLP_W2 + k * MLP_H + j];
    h2[j] = a > 0 ? a : 0;
  }
  for (int j = 0; j < MLP_OUT; ++j) {
    double a = w[MLP_B3 + j];
    for (int k = 0; k < MLP_H; ++k) a += h2[k] * w[MLP_W3 + k * MLP_OUT + j];
    y[j] = a;
  }
}

void orc_mlp_forward(const double *w, const double *x, int64_t B, double *y) {
  double h1[MLP_H], h2[MLP_H];
  for (int64_t i = 0; i < B; ++i)
    mlp_one(w, x + i * MLP_IN, h1, h2, y + i * MLP_OUT);
}

double orc_mlp_grad(const double *w, const double *x, const double *t,
                    const double *mask, int64_t B, double *grad) {
  double h1[MLP_H], h2[MLP_H], y[MLP_OUT], dy[MLP_OUT], dh2[MLP_H], dh1[MLP_H];
  double msum = 0, loss = 0;
  for (int64_t i = 0; i < B * MLP_OUT; ++i) msum += mask[i];
  const double scale = msum > 0 ? 1.0 / msum : 0.0;
  memset(grad, 0, sizeof(double) * MLP_NP);
  for (int64_t i = 0; i < B; ++i) {
    const double *xi = x + i * MLP_IN;
    mlp_one(w, xi, h1, h2, y);
    for (int c = 0; c < MLP_OUT; ++c) {
      const double mk = mask[i * MLP_OUT + c];
      const double d = y[c] - t[i * MLP_OUT + c];
      loss += mk * d * d;
      dy[c] = 2.0 * scale * mk * d;
      grad[MLP_B3 + c] += dy[c];
    }
    for (int k = 0; k < MLP_H; ++k) {
      double s = 0;
      for (int c = 0; c < MLP_OUT; ++c) {
        grad[MLP_W3 + k * MLP_OUT + c] += h2[k] * dy[c];
        s += w[MLP_W3 + k * MLP_OUT + c] * dy[c];
      }
      dh2[k] = h2[k] > 0 ? s : 0;
    }
    for (int j = 0; j < MLP_H; ++j) grad[MLP_B2 + j] += dh2[j];
    for (int k = 0; k < MLP_H; ++k) {
      double s = 0;
      for (int j = 0; j < MLP_H; ++j) {
        grad[MLP_W2 + k * MLP_H + j] += h1[k] * dh2[j];
        s += w[MLP_W2 + k * MLP_H + j] * dh2[j];
      }
      dh1[k] = h1[k] > 0 ? s : 0;
    }
    for (int j = 0; j < MLP_H; ++j) grad[MLP_B1 + j] += dh1[j];
    for (int k = 0; k < MLP_IN; ++k)
      for (int j = 0; j < MLP_H; ++j) grad[MLP_W1 + k * MLP_H + j] += xi[k] * dh1[j];
  }
  return loss * scale;
}
