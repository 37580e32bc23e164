// Driver around the reference's looper/union_find.h (built into oracle/_ref by
// oracle/Makefile; test infrastructure only).
//
// stdin:  n  npairs  then npairs lines "i0 i1"
// stdout: one line per node: "parent root weight" (parent = -1 and weight > 0
//         for roots), after looper::union_find::unify on every pair in order
//         (union_find.h:281-284, weighted union with path halving).
#include <looper/union_find.h>
#include <cstdio>
#include <vector>

int main() {
  int n = 0, m = 0;
  if (std::scanf("%d %d", &n, &m) != 2) return 1;
  std::vector<looper::union_find::node> nodes(n);
  for (int i = 0; i < m; ++i) {
    int a = 0, b = 0;
    if (std::scanf("%d %d", &a, &b) != 2) return 1;
    looper::union_find::unify(nodes, a, b);
  }
  for (int i = 0; i < n; ++i) {
    bool r = nodes[i].is_root();
    std::printf("%d %d %d\n", r ? -1 : nodes[i].parent(), looper::union_find::root_index(nodes, i),
                r ? nodes[i].weight() : 0);
  }
  return 0;
}
