/*
 * CPU oracle for the traversal engine — TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatements of GSWITCH's CPU validation routines, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker. Nothing in the product (gswitch_amd/) links, loads or calls it.
 *
 * Graph input everywhere: CSR with int64 row_ptr[n+1], int32 col[m]
 * (and uint32 w[m] where weighted). Adjacency lists may be in any order.
 *
 * Pinning (see tests/test_oracle.py): every function is checked against
 * closed forms (paths, stars, grids, complete graphs, disconnected pieces),
 * against independent textbook algorithms in this file (Dijkstra, Bellman-
 * Ford, union-find) and against scipy.sparse.csgraph on seeded graphs.
 */
#ifndef GSW_ORACLE_H
#define GSW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* BFS levels, -1 if unreached. Restates bfs_cpu (application/bfs.cu:38-62). */
void orc_bfs(int64_t n, const int64_t *row_ptr, const int32_t *col,
             int32_t root, int32_t *lvl);

/* SSSP by SPFA (queue-based Bellman-Ford), uint32 distances, UINT32_MAX if
 * unreached. Restates spfa (application/sssp.cu:55-89) with the reference's
 * int/INT_MAX widened to uint32/UINT32_MAX. */
void orc_spfa(int64_t n, const int64_t *row_ptr, const int32_t *col,
              const uint32_t *w, int32_t root, uint32_t *dist);

/* Independent checks for orc_spfa: binary-heap Dijkstra and n-1 round
 * Bellman-Ford (the latter only for tiny graphs). */
void orc_dijkstra(int64_t n, const int64_t *row_ptr, const int32_t *col,
                  const uint32_t *w, int32_t root, uint32_t *dist);
void orc_bellman_ford(int64_t n, const int64_t *row_ptr, const int32_t *col,
                      const uint32_t *w, int32_t root, uint32_t *dist);

/* Connected components, label = smallest vertex id of the component.
 * Restates CC_CPU (application/cc.cu:34-66): ids visited in ascending order,
 * each unlabelled id floods its component by BFS. */
void orc_cc(int64_t n, const int64_t *row_ptr, const int32_t *col,
            int32_t *label);

/* Independent check for orc_cc: union-find over the edge list with the
 * root chosen as the smaller id. */
/* Graph colouring rounds (application/gc.cu:15-98), colours from 1. */
void orc_gc(int64_t n, const int64_t *row_ptr, const int32_t *col,
            int32_t *color);
void orc_cc_unionfind(int64_t n, const int64_t *row_ptr, const int32_t *col,
                      int32_t *label);

/* PageRank, FP64 power iteration in GSWITCH's un-normalised form
 * (application/pagerank.cu:19-32, 119-125):
 *   pr_0[v]     = 1 - d
 *   pr_{k+1}[v] = (1 - d) + d * sum_{u in N(v)} pr_k[u] / outdeg(u)
 * run for `iters` steps; returns pr_iters. Vertices with outdeg 0
 * contribute nothing. */
void orc_pagerank(int64_t n, const int64_t *row_ptr, const int32_t *col,
                  double damping, int iters, double *pr);

/* Brandes betweenness from one root. Restates bc_cpu
 * (application/bc.cu:162-214): nsp = number of shortest paths (counted in
 * double to avoid the reference's int overflow), dep = dependency.
 * level[] receives the BFS levels (-1 unreached). */
void orc_bc(int64_t n, const int64_t *row_ptr, const int32_t *col,
            int32_t root, int32_t *level, double *nsp, double *dep);

/* 64-source reachability masks: bit s of mask[v] is set iff v is reachable
 * from sources[s]; nsrc <= 64. Defined as nsrc independent orc_bfs runs. */
void orc_msbfs64(int64_t n, const int64_t *row_ptr, const int32_t *col,
                 const int32_t *sources, int nsrc, uint64_t *mask);

/* Graph500 edge count of a traversal: (sum of deg(v) over v with
 * reached[v] != 0) / 2, for symmetric CSR. */
int64_t orc_traversed_edges(int64_t n, const int64_t *row_ptr,
                            const int32_t *reached_level);

/* Learned-selector network, FP64 (no reference counterpart: the reference
 * selects with trees only). 12 -> 64 -> 64 -> 30 with ReLU on both hidden
 * layers; w = W1[12][64] b1[64] W2[64][64] b2[64] W3[64][30] b3[30]
 * (6942 doubles, input index first). y[B][30]. */
void orc_mlp_forward(const double *w, const double *x, int64_t B, double *y);
/* L = sum(mask * (y - t)^2) / sum(mask) (0 if the mask is empty); grad
 * receives dL/dw (6942 doubles). Returns L. */
double orc_mlp_grad(const double *w, const double *x, const double *t,
                    const double *mask, int64_t B, double *grad);

#ifdef __cplusplus
}
#endif

#endif
