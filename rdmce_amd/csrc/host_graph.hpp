// Host graph: COO-text / .bin loading, vertex orderings and CSR.
//
// Mirrors the reference's `class Graph` (RDMCE/include/graph.h:12-52) but is
// stored as CSR from the start (u64 row offsets, u32 column ids) instead of
// vector<vector<u32>>, so the largest inputs (hundreds of millions of
// adjacency entries) are built with parallel radix sorts.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace rdmce {

enum class Order : int { UNCHANGED = 0, ASC = 1, DES = 2, DEG = 3 };

struct HostGraph {
  std::string name;
  uint64_t n = 0;            // vertices (every vertex has >= 1 edge)
  uint64_t m = 0;            // undirected edges
  uint64_t max_degree = 0;
  uint64_t degeneracy = 0;   // = max_degree until a degeneracy ordering is applied
  std::vector<uint64_t> rowoff;   // n + 1
  std::vector<uint32_t> col;      // 2m, each row sorted ascending
  std::vector<uint32_t> labels;   // n: original id of each vertex
  std::vector<uint32_t> core;     // n: core number (filled by the DEG ordering)

  uint64_t degree(uint64_t v) const { return rowoff[v + 1] - rowoff[v]; }
  void recompute_max_degree();
};

// Builds a graph from raw endpoint pairs with the reference loader's
// semantics (RDMCE/src/graph.cc:64-152): self loops dropped, pairs made
// canonical (min, max), sorted and deduplicated, and the ids that occur in a
// kept edge relabelled densely in ascending order (labels[new] = original).
void graph_from_pairs(HostGraph &g, const uint32_t *u, const uint32_t *v, uint64_t count);
// Same from packed canonical keys (min << 32 | max); consumes the vector.
void graph_from_keys(HostGraph &g, std::vector<uint64_t> &&keys);

// COO text loader (RDMCE/src/graph.cc:64-152).  Returns false when the file
// cannot be opened (the reference exits(1) there).
bool load_coo_text(HostGraph &g, const std::string &path, std::string *err);
// Parses COO text already in memory with the same rules.
void parse_coo_buffer(const char *buf, size_t len, std::vector<uint64_t> &keys);

// Reference .bin layout (RDMCE/src/graph.cc:154-213), little endian:
// u64 n, u64 m, u64 max_degree, u64 degeneracy; per vertex u64 degree then
// degree x u32 neighbour ids; then n x u32 labels.
bool load_bin(HostGraph &g, const std::string &path, std::string *err);
bool store_bin(const HostGraph &g, const std::string &path, std::string *err);
// Output name rule of StoreIntoBin (graph.cc:155-156): extension replaced by .bin.
std::string bin_path_for(const std::string &input_path);
// Base name without directory and extension (ExtractName, common.cu:2-17).
std::string extract_name(const std::string &path);

// Vertex orderings (Graph::SortByOrder, graph.cc:241-306).
void apply_order(HostGraph &g, Order order);
// Relabels so that new vertex i is old vertex order[i] (Graph::ApplyOrder,
// graph.cc:2-25).
void relabel(HostGraph &g, const std::vector<uint32_t> &order);
// The reference's degeneracy peel order (graph.cc:281-306), computed with
// degree buckets in O((n + m) log n) instead of the reference's O(n * delta)
// rescans; also fills core numbers.  Returns the degeneracy.
uint64_t degeneracy_order(const HostGraph &g, std::vector<uint32_t> &order,
                          std::vector<uint32_t> &core);

// Writes "u v" lines (original labels, u < v, one per undirected edge).
bool write_coo_text(const HostGraph &g, const std::string &path, std::string *err);

}  // namespace rdmce
