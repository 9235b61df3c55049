// Driver for the reference's own CPU code, built by oracle/Makefile into
// oracle/_ref/ from the reference sources where they lie.  TEST
// INFRASTRUCTURE ONLY: it pins this repository's loader, orderings and .bin
// writer (byte-exact .bin comparison) and its clique counts (the reference's
// vertex-rooted Tomita solver, RDMCE/src/mce.cc:64-197).
//
// usage: ref_mce <graph.txt> <a|d|deg|unchange> <out.bin> [count]
//   loads with Graph::LoadFromFile (graph.cc:64), orders with
//   Graph::SortByOrder (graph.cc:241), writes Graph::StoreIntoBin
//   (graph.cc:154; the path's extension is replaced by .bin), and with
//   "count" prints "count <n>" from MceSolver::Solve(BKPivot).
#include <cstdio>
#include <cstring>
#include "graph.h"
#include "mce.h"

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s graph.txt a|d|deg|unchange out.bin [count]\n", argv[0]);
    return 2;
  }
  OrderType o = OrderType::DEG;
  if (!std::strcmp(argv[2], "a")) o = OrderType::ASC;
  else if (!std::strcmp(argv[2], "d")) o = OrderType::DES;
  else if (!std::strcmp(argv[2], "unchange")) o = OrderType::UNCHANGED;
  Graph g;
  g.LoadFromFile(argv[1]);
  g.SortByOrder(o);
  g.StoreIntoBin(argv[3]);
  if (argc > 4 && !std::strcmp(argv[4], "count")) {
    MceSolver s(g);
    s.Solve(MceAlgorithm::BKPivot);
    std::printf("count %llu\n", (unsigned long long)s.mce_count_);
  }
  return 0;
}
