// Golden-table dumper for the reference's rule-based selector — TEST
// INFRASTRUCTURE ONLY. Compiled (oracle/Makefile, target `ref`) against the
// reference's own header-only decision trees in src/model/*.h, where they
// lie; the binary goes to oracle/_ref/ and is never committed.
//
// stdin : lines "<tree_name> <k> f0 f1 ... f(k-1)"
// stdout: one predicted class per line, in input order.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "model/select_dir_asso.h"
#include "model/select_dir_asso_backup.h"
#include "model/select_dir_idem.h"
#include "model/select_fmt.h"
#include "model/select_fusion.h"
#include "model/select_lb_asso_backward.h"
#include "model/select_lb_asso_backward_backup.h"
#include "model/select_lb_asso_forward.h"
#include "model/select_lb_asso_forward_backup.h"
#include "model/select_lb_idem.h"
#include "model/select_stepping.h"

typedef int (*tree_fn)(const std::vector<double> &);

static tree_fn lookup(const char *name) {
  struct {
    const char *n;
    tree_fn f;
  } table[] = {
      {"select_dir_asso", select_dir_asso},
      {"select_dir_asso_backup", select_dir_asso_backup},
      {"select_dir_idem", select_dir_idem},
      {"select_fmt", select_fmt},
      {"select_fusion", select_fusion},
      {"select_lb_asso_backward", select_lb_asso_backward},
      {"select_lb_asso_backward_backup", select_lb_asso_backward_backup},
      {"select_lb_asso_forward", select_lb_asso_forward},
      {"select_lb_asso_forward_backup", select_lb_asso_forward_backup},
      {"select_lb_idem", select_lb_idem},
      {"select_stepping", select_stepping},
  };
  for (auto &t : table)
    if (!strcmp(t.n, name)) return t.f;
  return nullptr;
}

int main() {
  char name[128];
  int k;
  while (scanf("%127s %d", name, &k) == 2) {
    std::vector<double> fv(k);
    for (int i = 0; i < k; ++i)
      if (scanf("%lf", &fv[i]) != 1) return 2;
    tree_fn f = lookup(name);
    if (!f) {
      fprintf(stderr, "unknown tree %s\n", name);
      return 3;
    }
    printf("%d\n", f(fv));
  }
  return 0;
}
