// Flat decision-tree evaluator for the rule-based selector.
// The reference evaluates its trees as generated if/else code over a
// std::vector<double> (src/model/select_*.h). Here a tree is a node table
// walked by one loop, so the same tree runs on the host and inside a kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace gsw_model {

struct tree_node_t {
  int feature;  // < 0: leaf
  int left;
  int right;
  int cls;
  double threshold;
};

struct tree_t {
  const tree_node_t *nodes;
  int n_nodes;
  int n_features;  // 1 + highest feature index read
};

__host__ __device__ inline int eval_tree(const tree_node_t *nodes,
                                         const double *fv) {
  int i = 0;
  while (nodes[i].feature >= 0)
    i = (fv[nodes[i].feature] <= nodes[i].threshold) ? nodes[i].left
                                                      : nodes[i].right;
  return nodes[i].cls;
}

inline int eval_tree(const tree_t &t, const double *fv) {
  return eval_tree(t.nodes, fv);
}

}  // namespace gsw_model
