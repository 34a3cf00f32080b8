/*
 * gswitch_amd — C-ABI of the MI355X traversal engine (libgswitch_amd.so).
 *
 * The reference exposes a header-only C++ template API and no C-ABI
 * (its apps #include "gswitch.h" and subclass Functor). This header is the
 * drop-in boundary for callers that bind over an FFI: plain pointers and
 * sizes, opaque handles with explicit destroy, a status code on every call
 * and a thread-local message from gsw_last_error(). Each entry point names
 * the reference interface it replaces (paths relative to the reference
 * tree). Host pointers are read/written synchronously; device work is
 * ordered on the library's own stream. Handles are thread-compatible (one
 * thread per handle), not thread-safe.
 */
#ifndef GSWITCH_AMD_H
#define GSWITCH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t gsw_status;
enum {
  GSW_OK = 0,
  GSW_E_INVALID = 1,     /* bad argument (null pointer, id out of range) */
  GSW_E_HIP = 2,         /* HIP runtime error, or no GPU present */
  GSW_E_ALLOC = 3,       /* device allocation failed */
  GSW_E_OVERFLOW = 4,    /* a frontier queue would overflow its capacity */
  GSW_E_UNSUPPORTED = 5, /* variant or option not available for this call */
  GSW_E_INTERNAL = 6
};

typedef struct gsw_host_graph gsw_host_graph; /* host CSR owned by the lib */
typedef struct gsw_graph gsw_graph;           /* device-resident graph     */

/* ---- library ---------------------------------------------------------- */
const char *gsw_version(void);
/* Message of the last failed call on this thread ("" if none). Replaces the
 * reference's print-and-exit error handling (H_ERR, src/utils/utils.cuh:70-78). */
const char *gsw_last_error(void);
/* Number of visible GPUs (0 without a GPU; no other call then succeeds). */
int32_t gsw_device_count(void);

/* ---- host ingest (src/utils/fileIO.cuh) -------------------------------
 * CSR build semantics of host_graph_t<CSR>::build + rinse
 * (src/utils/fileIO.cuh:210-332, 399-443): self-loops dropped, undirected
 * edges stored both ways, duplicates removed, adjacency sorted; vertex ids
 * are kept (no isolated-vertex renumbering). Weighted duplicates keep the
 * smallest weight. */
gsw_status gsw_host_graph_from_edges(int64_t n, int64_t ne, const int32_t *src,
                                     const int32_t *dst, const uint32_t *w,
                                     int32_t directed, gsw_host_graph **out);
/* Seeded synthetic generators (counter-based SplitMix64 draws). R-MAT:
 * edge_factor * 2^scale samples with quadrant probabilities (a, b, c,
 * 1-a-b-c). Weighted graphs get one uint32 weight in [1, 64] per undirected
 * edge, shared by both directions (as gen_weight, src/utils/fileIO.cuh:148-166,
 * shared at :288, :295). */
gsw_status gsw_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                        double c, uint64_t seed, int32_t weighted,
                        gsw_host_graph **out);
gsw_status gsw_gen_grid2d(int64_t rows, int64_t cols, double p_delete,
                          double long_range_frac, uint64_t seed,
                          int32_t weighted, gsw_host_graph **out);
/* Binary CSR cache, the reference's store_to_binfile / load_from_binfile
 * (src/utils/fileIO.cuh:334-387): save a host CSR (offsets, adjacency,
 * weights, reverse CSR of a directed graph) and load it back without
 * parsing the text file again. The layout is this library's own (tagged,
 * int64 offsets); a truncated or corrupt file is an error. */
gsw_status gsw_host_graph_save(const gsw_host_graph *hg, const char *path);
gsw_status gsw_host_graph_load(const char *path, gsw_host_graph **out);
/* Training records in the reference's format (feature_t::record,
 * src/abstraction/features.cuh:86-104, enabled there by --json=<path>):
 * from now on every expanded level of a vertex-centric run appends one JSON
 * object {"0": fv[0], ..., "20": fv[20], "21": last filter ms, "22": last
 * expand ms} to `path`, at most 100 per call. NULL or "" turns it off. */
gsw_status gsw_record_features(const char *path);
/* Matrix Market / edge-list file -> host CSR. Replaces
 * edgelist_t::read_mtx + host_graph_t<CSR>::build (+ rinse)
 * (src/utils/fileIO.cuh:73-146, 210-332, 399-443) behind build_graph<C,E>(path,
 * fets, with_header, with_weight, directed) (graph.cuh:268-280): comment lines
 * skipped, optional "rows cols nnz" header ignored, self-loops dropped, ids
 * shifted by the smallest id, a final record without a newline dropped,
 * integer weights when with_weight. rinse: 1 drops degree-zero vertices and
 * renumbers the rest (undirected only), 0 keeps every id, -1 does what the
 * reference does (rinse an undirected unweighted graph only). */
gsw_status gsw_read_mtx(const char *path, int32_t directed, int32_t with_weight,
                        int32_t with_header, int32_t rinse,
                        gsw_host_graph **out);
gsw_status gsw_gen_er(int64_t n, double mean_degree, uint64_t seed,
                      int32_t weighted, gsw_host_graph **out);
gsw_status gsw_gen_ba(int64_t n, int32_t m, uint64_t seed, int32_t weighted,
                      gsw_host_graph **out);
gsw_status gsw_host_graph_info(const gsw_host_graph *g, int64_t *n, int64_t *m,
                               int32_t *directed, int32_t *weighted);
/* Copy out row_ptr[n+1], col[m] and (if weighted and w != NULL) w[m]. */
gsw_status gsw_host_graph_export(const gsw_host_graph *g, int64_t *row_ptr,
                                 int32_t *col, uint32_t *w);
void gsw_host_graph_destroy(gsw_host_graph *g);
/* Degree statistics used as selector features: out[0..5] = avg, std, max,
 * range, Gini, entropy (graph_attr_t::fill, src/utils/fileIO.cuh:35-67). */
gsw_status gsw_degree_stats(const int64_t *row_ptr, int64_t n, double *out6);

/* ---- device graph (device_graph_t<CSR>, src/data_structures/graph.cuh:44-175;
 *      build_graph, graph.cuh:268-280) ----------------------------------- */
gsw_status gsw_graph_create_csr(const int64_t *row_ptr, const int32_t *col,
                                const uint32_t *w, int64_t n, int64_t m,
                                int32_t directed, gsw_graph **out);
gsw_status gsw_graph_from_host(const gsw_host_graph *hg, gsw_graph **out);
gsw_status gsw_graph_destroy(gsw_graph *g);

/* ---- runs ---------------------------------------------------------------
 * Options. `configs` pins the selector's choices like the reference's
 * --configs=Dir-Fmt-LB-Fusion (src/utils/cmdline.cuh:51-94); any field may
 * be "x"; Fmt also accepts "SortedQueue", LB also "TWC". */
typedef struct {
  const char *configs;
  int32_t max_levels; /* safety cap on loop iterations */
  int32_t verbose;    /* per-level log on stderr */
  int32_t twc_small;  /* TWC bins: degree < twc_small -> thread */
  int32_t twc_large;  /*           degree < twc_large -> wavefront */
  int32_t profile;    /* time every expand with hipEvents */
  int32_t pull_kernel; /* from-all pulls of unweighted graphs: 0 TWOD as
                          length-binned source slices, STRICT as a merge-path
                          walk; 1 both as segmented reductions; 2 (default)
                          TWOD by propagation blocking (LDS-sized source
                          chunks and target bins), STRICT as in 0 */
  int32_t capture;     /* gsw_bfs_multi: >0 = hipGraph-captured fused level loop,
                          2*capture levels per replay; 0 = per-level host loop */
  int32_t grid;        /* gsw_bfs_multi captured loop: workgroups per launch
                          (0 = sized from the graph) */
  int32_t selector;    /* 0 rule-based trees (src/model/select_*.h);
                          1 learned MLP (weights from gsw_selector_load);
                          2 class tree (gsw_selector_load_tree) */
} gsw_run_opts;
void gsw_run_opts_init(gsw_run_opts *o);

typedef struct {
  double device_ms;      /* hipEvent: frontier init -> results ready */
  int32_t levels;        /* level-loop iterations */
  int32_t n_records;     /* expands logged (gsw_last_records) */
  int64_t reached;       /* vertices with a result (BFS/SSSP), components (CC) */
} gsw_run_stats;

/* One record per expand launch (the reference's --verbose line,
 * src/abstraction/executor.cuh:138-146, plus counters). */
typedef struct {
  int32_t level;
  int8_t dir;    /* 0 Push, 1 Pull */
  int8_t fmt;    /* 0 Queue, 1 Bitmap */
  int8_t lb;     /* 0 TM, 1 WM, 2 CM, 3 STRICT, 4 TWOD, 5 ELB, 6 TWC */
  int8_t fused;  /* 1 if inspect+filter were fused into the expand */
  int8_t sorted; /* queue format: 1 sorted, 0 unsorted */
  int8_t kernel; /* pulls: the from-all pull kernel in effect (gsw_run_opts.pull_kernel) */
  int8_t pad[2];
  int64_t frontier; /* input frontier length */
  int64_t active;   /* active vertices (edges for CC) at this level */
  int64_t inactive;
  int64_t edges;    /* edges inspected by the expand */
  int64_t verts;    /* frontier vertices expanded */
  float ms;         /* expand duration (0 unless profile) */
  float pad2;
} gsw_level_record;

/* BFS levels (-1 unreached) from `root`. Replaces run_bfs
 * (application/bfs.cu:78-102) with the BFS functor of bfs.cu:16-36. */
gsw_status gsw_bfs(gsw_graph *g, int32_t root, const gsw_run_opts *opts,
                   int32_t *levels_out, gsw_run_stats *stats);
/* SSSP distances (UINT32_MAX unreached); needs a weighted graph. Replaces
 * run_sssp (application/sssp.cu:106-126), functor sssp.cu:17-53, stepping
 * window src/data_structures/window.cuh:44-123. */
gsw_status gsw_sssp(gsw_graph *g, int32_t root, const gsw_run_opts *opts,
                    uint32_t *dist_out, gsw_run_stats *stats);
/* Connected components, label = smallest vertex id in the component
 * (run_cc, application/cc.cu:82-108). By default: tree hooking over the CSR
 * with subgraph sampling (two sampled neighbours per vertex, then the
 * remaining edges of every vertex outside the most frequent sampled
 * component). With opts->configs set: the reference's edge-centric hooking
 * (cc.cu:15-32) + pointer jumping (src/utils/intrinsics.cuh:328-362) through
 * the level loop, honouring the pinned choices. With opts->selector 1 or 2
 * and no configs, the learned network or class tree of GSW_APP_CC chooses
 * between sampled hooking and the edge-centric loop (rule, bitmap or queue
 * frontiers) once per run (gsw_last_class). */
gsw_status gsw_cc(gsw_graph *g, const gsw_run_opts *opts, int32_t *labels_out,
                  gsw_run_stats *stats);
/* Allocation-free form of gsw_cc's default path (sampled hooking): no
 * counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops). `labels` is DEVICE
 * memory of n int32 owned by the caller and receives the labels;
 * `stream` is a hipStream_t (NULL: the null stream). The call only queues
 * kernels and returns; it allocates nothing, reads nothing back and
 * synchronises nothing. `samples` (1..4096; gsw_cc uses 1024) vertices
 * are sampled to pick the component whose inner edges are skipped. The
 * sample and the chosen root live in one static device array per GPU
 * (GSW_CC_NOALLOC_SCRATCH_BYTES, held until the process ends) that every
 * call on that GPU shares: calls that overlap on one GPU race. */
#define GSW_CC_NOALLOC_MAX_SAMPLES 4096
#define GSW_CC_NOALLOC_SCRATCH_BYTES (4 * (4096 + 64))
gsw_status gsw_cc_noalloc(const gsw_graph *g, int32_t samples, int32_t *labels,
                          void *stream);
/* PageRank, un-normalised: pr_0 = 1-d, pr_{k+1}(v) = 1-d + d*sum pr_k(u)/deg(u),
 * `iters` iterations; eps > 0 additionally deactivates vertices whose
 * relative change is below eps as the reference does. Replaces
 * run_pagerank (application/pagerank.cu:81-103), functor pagerank.cu:18-42.
 * Every iteration pulls every vertex the same way, so the variant is chosen
 * once per run: by the rule trees (selector 0), or among the 8 from-all
 * pull classes by the learned network or class tree of GSW_APP_PR
 * (selector 1 / 2; gsw_last_class). */
gsw_status gsw_pagerank(gsw_graph *g, float damping, int32_t iters, float eps,
                        const gsw_run_opts *opts, float *pr_out,
                        gsw_run_stats *stats);
/* Delta-PageRank, described in the reference's documentation
 * (docs/index.md:175-179) but absent from its code: vertices send only the
 * change of their value; an accumulation of incoming changes is applied
 * (and sent on) once it exceeds eps. Converges to the fixed point of
 * gsw_pagerank's iteration, value = (1 - d) + d * sum value[u] / deg[u];
 * stops when no accumulation exceeds eps or after max_iters rounds. */
gsw_status gsw_delta_pagerank(gsw_graph *g, float damping, double eps,
                              int32_t max_iters, const gsw_run_opts *opts,
                              float *pr_out, gsw_run_stats *stats);
/* Graph colouring, run_gc (application/gc.cu:15-127): Jones-Plassmann rounds
 * with vertex ids as priorities; colors_out[v] = the round (from 1) in which
 * v's id first exceeded every uncoloured neighbour's. */
gsw_status gsw_gc(gsw_graph *g, const gsw_run_opts *opts, int32_t *colors_out,
                  gsw_run_stats *stats);
/* Betweenness dependencies from one root (Brandes): levels_out = BFS level
 * (-1 unreached), nsp_out = number of shortest root->v paths, dep_out =
 * dependency of the root on v. Any output may be null. Replaces run_bc
 * (application/bc.cu:97-160) with functors bc.cu:16-94; nsp and dep are
 * FP64 where the reference keeps int and float. Undirected graphs only. */
gsw_status gsw_bc(gsw_graph *g, int32_t root, const gsw_run_opts *opts,
                  int32_t *levels_out, double *nsp_out, double *dep_out,
                  gsw_run_stats *stats);
/* 64-source bitmask BFS: bit s of mask_out[v] is set iff v is reachable from
 * sources[s] (1 <= nsrc <= 64). level_sums_out[s] (64 entries, optional) =
 * sum over reached v of the BFS level of v from sources[s]. No counterpart
 * in the reference, whose application/mbfs.cu:122-153 re-runs the
 * single-source BFS; results equal nsrc independent gsw_bfs runs. */
gsw_status gsw_msbfs64(gsw_graph *g, const int32_t *sources, int32_t nsrc,
                       const gsw_run_opts *opts, uint64_t *mask_out,
                       uint64_t *level_sums_out, gsw_run_stats *stats);
/* Multi-root BFS: levels from the nearest of `roots` (duplicates ignored).
 * On a disjoint union of graphs (gsw_gen_er_batch) with one root per
 * instance this is one batched BFS over all instances: each instance's
 * levels equal its own single-source BFS (application/bfs.cu:78-102).
 * opts->capture > 0 runs a hipGraph-captured fused push loop (no host sync
 * inside a replay) with the LB of opts->configs (WM or TWC). */
gsw_status gsw_bfs_multi(gsw_graph *g, const int32_t *roots, int32_t nroots,
                         const gsw_run_opts *opts, int32_t *levels_out,
                         gsw_run_stats *stats);
/* Batched BFS + CC over a disjoint union of `count` instances (instance i =
 * vertices offsets[i] .. offsets[i+1]-1, offsets[0] = 0, offsets[count] = n,
 * as gsw_gen_er_batch lays them out): BFS levels from every instance's
 * first vertex, by the device-resident direction-optimising loop from all
 * the roots at once (the direction left free; opts->capture levels per
 * replay, opts->grid), or with a forced "Push-..." by the captured fused
 * push loop (opts->configs LB WM or TWC, opts->capture, opts->grid, TWC
 * thresholds); then connected components. When no edge leaves an
 * instance (checked once per offsets: rows are sorted, so each row's first
 * and last column decide), the vertices the BFS reached take their
 * instance's first vertex as label and only the others are hooked;
 * otherwise sampled hooking runs as one captured launch sequence, skipping
 * in each instance the edges inside the component of its first vertex
 * (edges that leave an instance are always linked). Each instance's levels
 * equal its own run_bfs from its first vertex (application/bfs.cu:78-102)
 * and its labels its own run_cc (application/cc.cu:82-108). device_ms covers
 * both. Undirected graphs only. No counterpart in the reference. */
gsw_status gsw_batch_bfs_cc(gsw_graph *g, const int64_t *offsets, int32_t count,
                            const gsw_run_opts *opts, int32_t *levels_out,
                            int32_t *labels_out, gsw_run_stats *stats);
/* SSSP stepping window of a graph: gsw_sssp's initial window is the
 * reference's init_conf size (8 * mean weight * 32 / mean degree,
 * src/gswitch.h:19-63) times `scale` (default 1). The distances do not depend
 * on it; the work per level does (a wide window re-relaxes vertices whose
 * distance is not final yet, a narrow one runs more levels). No counterpart
 * in the reference. */
gsw_status gsw_sssp_set_window(gsw_graph *g, double scale);
gsw_status gsw_sssp_get_window(const gsw_graph *g, double *scale);
/* Autotuner for the window: times gsw_sssp from `roots` (best of 2 each,
 * opts as given, e.g. the learned selector) under the scales 1, 1/2, 1/4,
 * 1/8 and 1/16, keeps the fastest total as the graph's window scale and
 * returns it. */
gsw_status gsw_autotune_sssp(gsw_graph *g, const int32_t *roots, int32_t nroots,
                             const gsw_run_opts *opts, double *best_scale,
                             double *best_ms);
/* Autotuner for the captured batched BFS: times {WM, TWC} x TWC bin
 * thresholds x workgroups per launch on this graph and roots (best of 3)
 * and returns the fastest (lb uses the gsw_level_record codes; feed the
 * result back through gsw_run_opts configs/twc_small/twc_large/grid).
 * No counterpart in the reference, whose grid sizes and thresholds are
 * compile-time constants (src/utils/common.cuh:14-24). */
gsw_status gsw_autotune_bfs(gsw_graph *g, const int32_t *roots, int32_t nroots,
                            int32_t capture, int32_t *best_lb,
                            int32_t *best_small, int32_t *best_large,
                            int32_t *best_grid, double *best_ms);
/* Batch of Erdos-Renyi graphs as one disjoint union: instance i has
 * floor(exp(U[ln nmin, ln nmax])) vertices and mean degree U[dmin, dmax]
 * (both drawn from the batch seed), its own edge stream, and vertex ids
 * offsets_out[i] .. offsets_out[i+1]-1 (count + 1 entries). */
gsw_status gsw_gen_er_batch(int32_t count, int64_t nmin, int64_t nmax,
                            double dmin, double dmax, uint64_t seed,
                            int32_t weighted, gsw_host_graph **out,
                            int64_t *offsets_out);
/* Instances first .. first+count-1 of the same batch, numbered from vertex
 * 0 (offsets_out[0] = 0): consecutive ranges cut the whole batch into
 * pieces, so a batch too large for one union is built and run piecewise. */
gsw_status gsw_gen_er_batch_range(int64_t first, int32_t count, int64_t nmin,
                                  int64_t nmax, double dmin, double dmax,
                                  uint64_t seed, int32_t weighted,
                                  gsw_host_graph **out, int64_t *offsets_out);
/* Records of the last run on `g`; *n = number copied (<= cap). */
gsw_status gsw_last_records(const gsw_graph *g, gsw_level_record *out,
                            int32_t cap, int32_t *n);

/* ---- rule-based selector (src/model/select_*.h) ------------------------
 * Evaluate one of the reference's trained decision trees by name
 * ("select_dir_idem", "select_lb_idem", ...) on a feature vector. */
gsw_status gsw_tree_eval(const char *name, const double *fv, int32_t nfv,
                         int32_t *cls);

/* ---- learned selector (no counterpart in the reference, which trains its
 *      trees offline from feature dumps: src/abstraction/features.cuh:86-104)
 * MLP 12 -> 64 -> 64 -> 30, FP32, 6942 parameters laid out as
 * W1[12][64] b1[64] W2[64][64] b2[64] W3[64][30] b3[30]. Output c predicts
 * log2(cost of variant class c / cost of the fastest class) for a level;
 * class = group*5 + lb, lb in {TM, WM, CM, STRICT, TWC}, groups
 * Push-Queue, Push-SortedQueue, Push-Bitmap, Push-Queue-Fused, Pull-Queue,
 * Pull-Bitmap. Inputs: see gsw_mlp_features. */
#define GSW_MLP_NPARAM 6942
#define GSW_MLP_NIN 12
#define GSW_MLP_NOUT 30
/* Applications with a learned selector of their own. */
#define GSW_APP_BFS 0
#define GSW_APP_SSSP 1
#define GSW_APP_PR 2  /* classes: the from-all pull kernels, chosen once per run */
#define GSW_APP_CC 3  /* classes: sampled hooking or the edge-centric loop, once per run */
#define GSW_APPS 4
/* Install the weights that `app`'s runs use when gsw_run_opts.selector == 1. */
gsw_status gsw_selector_load_app(int32_t app, const float *w, int32_t nparam);
/* The class the learned selector (selector 1 or 2) chose in the graph's
 * latest PageRank or CC run, applications that choose once per run
 * (PageRank: 0 TM, 1 WM, 2 CM, 3 STRICT merge-path walk, 4 STRICT
 * segmented, 5 TWOD length-binned slices, 6 TWOD segmented slices, 7 TWOD
 * propagation blocking; CC: 0 sampled hooking, 1 edge-centric with the
 * rule's formats, 2 edge-centric bitmap, 3 edge-centric queue); -1 if that
 * run used no learned selector (or a class tree's leaf it cannot run).
 * No counterpart in the reference. */
gsw_status gsw_last_class(const gsw_graph *g, int32_t *cls);
/* = gsw_selector_load_app(GSW_APP_BFS, w, nparam). */
gsw_status gsw_selector_load(const float *w, int32_t nparam);
/* Install the class tree that `app`'s runs use when gsw_run_opts.selector
 * == 2: the reference's offline-trained decision trees (src/model/select_*.h,
 * trained from feature_t::record dumps, README.md:13) retrained for this
 * engine, mapping the 12 inputs of gsw_mlp_features straight to a variant
 * class. Node i is internal if feat[i] >= 0 (x[feat[i]] <= thr[i] -> left[i],
 * else right[i]; children have larger indices) and a leaf of class cls[i]
 * otherwise; node 0 is the root. A leaf class the application cannot run
 * falls back to the rule trees for that level. */
gsw_status gsw_selector_load_tree(int32_t app, int32_t nodes, const int32_t *feat,
                                  const float *thr, const int32_t *left,
                                  const int32_t *right, const int32_t *cls);
/* The 12 inputs from counts[11] = {n, m, avg deg, std deg, Gini, level,
 * active, inactive, push workload, pull workload, max active degree}:
 * log2(1+n)/24, log2(1+m)/32, log2(1+avg)/8, log2(1+std/avg)/4, Gini,
 * log2(1+level)/8, active/n, inactive/n, log2(1+push)/32, log2(1+pull)/32,
 * log2(1+push/max(active,1))/16, log2(1+max active degree)/24. */
gsw_status gsw_mlp_features(const double *counts, float *x);
/* Device forward pass: y[B][30]. */
gsw_status gsw_mlp_forward(const float *w, const float *x, int64_t B, float *y);
/* Device selection (the inference path of the inspector): cls[B] = argmin of
 * the outputs over the classes set in `mask`. */
gsw_status gsw_mlp_select(const float *w, const float *x, int64_t B,
                          uint32_t mask, int32_t *cls);
/* Device loss and gradient of L = sum(mask*(y-t)^2) / sum(mask). */
gsw_status gsw_mlp_grad(const float *w, const float *x, const float *t,
                        const float *mask, int64_t B, float *loss,
                        float *grad);
/* Adam training on the device (minibatches in a seeded shuffled order,
 * deterministic reductions): w_inout holds the initial weights and receives
 * the trained ones; loss_curve (optional) gets the mean loss per epoch. */
gsw_status gsw_mlp_train(const float *x, const float *t, const float *mask,
                         int64_t B, int32_t epochs, int32_t batch, float lr,
                         uint64_t seed, float *w_inout, float *loss_curve);
/* Exhaustive per-level timing for BFS from `root`: for each probed level
 * (every `stride`-th, at most max_levels) the 12 features, the wall time in
 * ms of each of the 30 classes run from a snapshot of the level state
 * (best of `reps`; a standalone class pays its filter, expand and the
 * inspection of the next level, a fused one its expand and the read of the
 * next queue length), and the class the rule selector would pick.
 * *n_out = levels written. */
gsw_status gsw_explore_bfs(gsw_graph *g, int32_t root, int32_t max_levels,
                           int32_t stride, int32_t reps, float *feat_out,
                           float *ms_out, int32_t *rule_out, int32_t *level_out,
                           int32_t *n_out);
/* The same for SSSP from `root` on a weighted graph: the state carried from
 * level to level is the two distance arrays, the stepping window and the
 * features; levels where only the window advances are not probed. */
gsw_status gsw_explore_sssp(gsw_graph *g, int32_t root, int32_t max_levels,
                           int32_t stride, int32_t reps, float *feat_out,
                           float *ms_out, int32_t *rule_out, int32_t *level_out,
                           int32_t *n_out);

/* ---- primitives (replace moderngpu's mgpu::scan / sorted_search used at
 *      src/kernel_libs/expand_VC_STRICT.cuh:540-552) --------------------- */
/* out[i] = sum(in[0..i)), out[n] = total; host buffers, device scan. */
gsw_status gsw_scan_exclusive_i64(const int64_t *in, int64_t n, int64_t *out);
/* wave64 DPP inclusive scan of 64 ints (host buffers); diagnostics. */
gsw_status gsw_wave_scan64(const int32_t *in, int32_t *out);
/* Device-to-device copy of `bytes` with a plain 16-byte-per-lane kernel,
 * `reps` times; *gbps = best (read + write) bandwidth. The measured HBM
 * ceiling, and the calibration kernel for hardware byte counters; it has
 * no counterpart in the reference. */
gsw_status gsw_stream_copy(int64_t bytes, int32_t reps, double *gbps);

#ifdef __cplusplus
}
#endif

#endif /* GSWITCH_AMD_H */
