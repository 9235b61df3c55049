#include "host_graph.hpp"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <fstream>

#include "parallel.hpp"

namespace rdmce {

void HostGraph::recompute_max_degree() {
  max_degree = 0;
  for (uint64_t v = 0; v < n; ++v) max_degree = std::max(max_degree, degree(v));
}

// ---------------------------------------------------------------------------
// COO text parsing.  The reference reads with `fin >> u >> v` into uint32_t
// and, on success or failure alike, discards the rest of the line
// (graph.cc:83-104).  libstdc++'s unsigned extraction: leading whitespace
// (newlines included) is skipped, an optional sign is accepted, a '-' value is
// negated modulo 2^32, a magnitude above 2^32-1 fails, and parsing stops at
// the first non-digit.
// ---------------------------------------------------------------------------
static inline bool is_space(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

static inline bool read_u32(const char *b, size_t len, size_t &pos, uint32_t &out) {
  while (pos < len && is_space(b[pos])) ++pos;
  if (pos >= len) return false;
  bool neg = false;
  if (b[pos] == '+' || b[pos] == '-') {
    neg = b[pos] == '-';
    ++pos;
  }
  if (pos >= len || b[pos] < '0' || b[pos] > '9') return false;
  uint64_t val = 0;
  bool over = false;
  while (pos < len && b[pos] >= '0' && b[pos] <= '9') {
    if (!over) {
      val = val * 10 + (uint64_t)(b[pos] - '0');
      if (val > 0xFFFFFFFFull) over = true;
    }
    ++pos;
  }
  if (over) return false;
  out = neg ? (uint32_t)(0u - (uint32_t)val) : (uint32_t)val;
  return true;
}

void parse_coo_buffer(const char *buf, size_t len, std::vector<uint64_t> &keys) {
  size_t pos = 0;
  while (pos < len) {
    uint32_t u, v;
    bool ok = read_u32(buf, len, pos, u) && read_u32(buf, len, pos, v);
    if (ok && u != v) {
      if (u > v) std::swap(u, v);
      keys.push_back(((uint64_t)u << 32) | v);
    }
    while (pos < len && buf[pos] != '\n') ++pos;
    if (pos < len) ++pos;
  }
}

static int bits_for(uint64_t maxval) {
  int b = 1;
  while (b < 64 && (maxval >> b) != 0) ++b;
  return b;
}

void graph_from_keys(HostGraph &g, std::vector<uint64_t> &&keys) {
  // Sort + unique canonical pairs (graph.cc:109-111).
  uint64_t max_label = 0;
  for (uint64_t k : keys) max_label = std::max<uint64_t>(max_label, k & 0xFFFFFFFFull);
  radix_sort_u64(keys, 32 + bits_for(max_label));
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  const uint64_t m = keys.size();

  // Dense relabel of used ids in ascending order (graph.cc:123-130), via a
  // presence bitmap and per-word ranks (memory bounded by the label range).
  const uint64_t nwords = m ? (max_label >> 6) + 1 : 0;
  std::vector<uint64_t> present(nwords, 0);
  for (uint64_t k : keys) {
    uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
    present[a >> 6] |= 1ull << (a & 63);
    present[b >> 6] |= 1ull << (b & 63);
  }
  std::vector<uint32_t> rank(nwords + 1, 0);
  for (uint64_t w = 0; w < nwords; ++w) rank[w + 1] = rank[w] + (uint32_t)__builtin_popcountll(present[w]);
  const uint64_t n = nwords ? rank[nwords] : 0;
  auto to_id = [&](uint32_t x) -> uint32_t {
    uint64_t w = x >> 6;
    uint64_t below = present[w] & ((1ull << (x & 63)) - 1);
    return rank[w] + (uint32_t)__builtin_popcountll(below);
  };

  g.n = n;
  g.m = m;
  g.labels.assign(n, 0);
  for (uint64_t w = 0; w < nwords; ++w) {
    uint64_t bits = present[w];
    uint32_t r = rank[w];
    while (bits) {
      int b = __builtin_ctzll(bits);
      g.labels[r++] = (uint32_t)(w * 64 + b);
      bits &= bits - 1;
    }
  }

  // Both directions of every edge, sorted by (row, col), give the CSR with
  // every row ascending - the same adjacency the reference builds by pushing
  // (graph.cc:135-145).
  std::vector<uint64_t> dkeys(2 * m);
  parallel_chunks(m, [&](int, uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      uint64_t a = to_id((uint32_t)(keys[i] >> 32)), c = to_id((uint32_t)keys[i]);
      dkeys[2 * i] = (a << 32) | c;
      dkeys[2 * i + 1] = (c << 32) | a;
    }
  });
  std::vector<uint64_t>().swap(keys);
  std::vector<uint64_t>().swap(present);
  std::vector<uint32_t>().swap(rank);
  radix_sort_u64(dkeys, 32 + bits_for(n ? n - 1 : 0));

  g.rowoff.assign(n + 1, 0);
  g.col.resize(2 * m);
  parallel_chunks(2 * m, [&](int, uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      g.col[i] = (uint32_t)dkeys[i];
      uint64_t r = dkeys[i] >> 32;
      uint64_t prev = i ? (dkeys[i - 1] >> 32) + 1 : 0;
      if (i == 0 || (dkeys[i - 1] >> 32) != r)
        for (uint64_t x = prev; x <= r; ++x) g.rowoff[x] = i;
    }
  });
  g.rowoff[n] = 2 * m;
  g.core.clear();
  g.recompute_max_degree();
  g.degeneracy = g.max_degree;  // graph.cc:150 (until a DEG ordering)
}

void graph_from_pairs(HostGraph &g, const uint32_t *u, const uint32_t *v, uint64_t count) {
  std::vector<uint64_t> keys;
  keys.reserve(count);
  for (uint64_t i = 0; i < count; ++i) {
    uint32_t a = u[i], b = v[i];
    if (a == b) continue;
    if (a > b) std::swap(a, b);
    keys.push_back(((uint64_t)a << 32) | b);
  }
  graph_from_keys(g, std::move(keys));
}

std::string extract_name(const std::string &path) {
  auto slash = path.find_last_of("/\\");
  std::string f = slash == std::string::npos ? path : path.substr(slash + 1);
  auto dot = f.find_last_of('.');
  return dot == std::string::npos ? f : f.substr(0, dot);
}

bool load_coo_text(HostGraph &g, const std::string &path, std::string *err) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) {
    if (err) *err = "Failed to open file " + path;
    return false;
  }
  std::vector<char> buf;
  std::vector<uint64_t> keys;
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz > 0) {
    buf.resize((size_t)sz);
    size_t got = std::fread(buf.data(), 1, (size_t)sz, f);
    buf.resize(got);
  }
  std::fclose(f);
  parse_coo_buffer(buf.data(), buf.size(), keys);
  std::vector<char>().swap(buf);
  graph_from_keys(g, std::move(keys));
  g.name = extract_name(path);
  return true;
}

std::string bin_path_for(const std::string &path) {
  size_t dot = path.find_last_of('.');
  return path.substr(0, dot) + ".bin";
}

bool store_bin(const HostGraph &g, const std::string &path, std::string *err) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) {
    if (err) *err = "Failed to open file " + path;
    return false;
  }
  uint64_t hdr[4] = {g.n, g.m, g.max_degree, g.degeneracy};
  bool ok = std::fwrite(hdr, sizeof(hdr), 1, f) == 1;
  for (uint64_t v = 0; v < g.n && ok; ++v) {
    uint64_t d = g.degree(v);
    ok = std::fwrite(&d, sizeof(d), 1, f) == 1;
    if (ok && d) ok = std::fwrite(&g.col[g.rowoff[v]], sizeof(uint32_t), d, f) == d;
  }
  if (ok && g.n) ok = std::fwrite(g.labels.data(), sizeof(uint32_t), g.n, f) == g.n;
  if (std::fclose(f) != 0) ok = false;
  if (!ok && err) *err = "write failed for " + path;
  return ok;
}

bool load_bin(HostGraph &g, const std::string &path, std::string *err) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) {
    if (err) *err = "Failed to open file " + path;
    return false;
  }
  auto fail = [&](const char *what) {
    std::fclose(f);
    if (err) *err = std::string(what) + " in " + path;
    return false;
  };
  uint64_t hdr[4];
  if (std::fread(hdr, sizeof(hdr), 1, f) != 1) return fail("truncated header");
  g.n = hdr[0];
  g.m = hdr[1];
  g.max_degree = hdr[2];
  g.degeneracy = hdr[3];
  if (g.n > (1ull << 32) || g.m > (1ull << 40)) return fail("implausible header");
  g.rowoff.assign(g.n + 1, 0);
  g.col.clear();
  g.col.reserve(2 * g.m);
  for (uint64_t v = 0; v < g.n; ++v) {
    uint64_t d;
    if (std::fread(&d, sizeof(d), 1, f) != 1) return fail("truncated adjacency");
    if (d > g.n) return fail("bad degree");
    size_t at = g.col.size();
    g.col.resize(at + d);
    if (d && std::fread(&g.col[at], sizeof(uint32_t), d, f) != d) return fail("truncated adjacency");
    g.rowoff[v + 1] = g.col.size();
  }
  g.labels.assign(g.n, 0);
  if (g.n && std::fread(g.labels.data(), sizeof(uint32_t), g.n, f) != g.n) return fail("truncated labels");
  std::fclose(f);
  if (g.col.size() != 2 * g.m) {
    if (err) *err = "edge count mismatch in " + path;
    return false;
  }
  for (uint32_t c : g.col)
    if (c >= g.n) {
      if (err) *err = "neighbour id out of range in " + path;
      return false;
    }
  g.core.clear();
  g.name = extract_name(path);
  return true;
}

bool write_coo_text(const HostGraph &g, const std::string &path, std::string *err) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) {
    if (err) *err = "Failed to open file " + path;
    return false;
  }
  for (uint64_t v = 0; v < g.n; ++v)
    for (uint64_t i = g.rowoff[v]; i < g.rowoff[v + 1]; ++i)
      if (g.col[i] > v) std::fprintf(f, "%u %u\n", g.labels[v], g.labels[g.col[i]]);
  return std::fclose(f) == 0;
}

// ---------------------------------------------------------------------------
// Orderings
// ---------------------------------------------------------------------------
void relabel(HostGraph &g, const std::vector<uint32_t> &order) {
  const uint64_t n = g.n;
  std::vector<uint32_t> rev(n);
  for (uint64_t i = 0; i < n; ++i) rev[order[i]] = (uint32_t)i;
  std::vector<uint64_t> off(n + 1, 0);
  for (uint64_t i = 0; i < n; ++i) off[i + 1] = off[i] + g.degree(order[i]);
  std::vector<uint32_t> col(g.col.size());
  parallel_for_dynamic(n, 1024, [&](uint64_t i) {
    uint32_t old = order[i];
    uint32_t *dst = &col[off[i]];
    const uint32_t *src = &g.col[g.rowoff[old]];
    uint64_t d = g.degree(old);
    for (uint64_t k = 0; k < d; ++k) dst[k] = rev[src[k]];
    std::sort(dst, dst + d);
  });
  std::vector<uint32_t> labels(n);
  for (uint64_t i = 0; i < n; ++i) labels[i] = g.labels[order[i]];
  if (!g.core.empty()) {
    std::vector<uint32_t> core(n);
    for (uint64_t i = 0; i < n; ++i) core[i] = g.core[order[i]];
    g.core.swap(core);
  }
  g.rowoff.swap(off);
  g.col.swap(col);
  g.labels.swap(labels);
}

uint64_t degeneracy_order(const HostGraph &g, std::vector<uint32_t> &order, std::vector<uint32_t> &core) {
  // Same visiting order as the reference peel: at level k, first every
  // unremoved vertex whose remaining degree equals k in ascending id, then
  // a FIFO pass in which each removed vertex decrements its neighbours (in
  // adjacency order) and a neighbour reaching exactly k is appended.
  // Instead of rescanning all n vertices per level, vertices are filed into
  // the bucket of each remaining degree they take above the current level.
  const uint64_t n = g.n;
  order.clear();
  order.reserve(n);
  core.assign(n, 0);
  std::vector<uint32_t> deg(n);
  uint64_t maxd = 0;
  for (uint64_t v = 0; v < n; ++v) {
    deg[v] = (uint32_t)g.degree(v);
    maxd = std::max<uint64_t>(maxd, deg[v]);
  }
  std::vector<std::vector<uint32_t>> bucket(maxd + 2);
  for (uint64_t v = 0; v < n; ++v) bucket[deg[v]].push_back((uint32_t)v);
  std::vector<uint8_t> removed(n, 0);
  std::vector<uint32_t> cand;
  uint64_t level = 0;
  while (order.size() < n) {
    ++level;
    size_t cur = order.size();
    if (level < bucket.size()) {
      cand.clear();
      for (uint32_t v : bucket[level])
        if (!removed[v] && deg[v] == level) cand.push_back(v);
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      for (uint32_t v : cand) {
        removed[v] = 1;
        core[v] = (uint32_t)level;
        order.push_back(v);
      }
      std::vector<uint32_t>().swap(bucket[level]);
    }
    while (cur < order.size()) {
      uint32_t v = order[cur++];
      for (uint64_t i = g.rowoff[v]; i < g.rowoff[v + 1]; ++i) {
        uint32_t u = g.col[i];
        uint32_t d = --deg[u];
        if (removed[u]) continue;
        if (d == level) {
          removed[u] = 1;
          core[u] = (uint32_t)level;
          order.push_back(u);
        } else if (d > level) {
          bucket[d].push_back(u);
        }
      }
    }
  }
  return level;
}

void apply_order(HostGraph &g, Order order) {
  if (order == Order::UNCHANGED) return;
  std::vector<uint32_t> ord(g.n);
  if (order == Order::DEG) {
    std::vector<uint32_t> core;
    uint64_t d = degeneracy_order(g, ord, core);
    g.core.swap(core);
    relabel(g, ord);
    g.degeneracy = d;
    return;
  }
  for (uint64_t v = 0; v < g.n; ++v) ord[v] = (uint32_t)v;
  // Same comparison as SortByDegreeAsc/Des (graph.cc:259-279), with
  // std::sort, so that ties land where the reference puts them.
  if (order == Order::ASC)
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return g.degree(a) < g.degree(b); });
  else
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return g.degree(a) > g.degree(b); });
  g.core.clear();
  relabel(g, ord);
}

}  // namespace rdmce
