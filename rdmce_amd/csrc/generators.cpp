#include "generators.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>

#include "parallel.hpp"
#include "rng.hpp"

namespace rdmce {

void gen_er(uint32_t n, double p, uint64_t seed, std::vector<uint64_t> &keys) {
  Xoshiro256ss rng(seed);
  for (uint32_t u = 0; u < n; ++u)
    for (uint32_t v = u + 1; v < n; ++v)
      if (rng.next_double() < p) keys.push_back(((uint64_t)u << 32) | v);
}

void gen_rmat(int scale, uint64_t edge_factor, double a, double b, double c, uint64_t seed,
              std::vector<uint64_t> &keys) {
  const uint64_t samples = edge_factor << scale;
  const uint64_t kBlock = 1ull << 20;
  const uint64_t nblocks = (samples + kBlock - 1) / kBlock;
  auto thr = [](double x) -> uint64_t {
    double t = std::ldexp(x, 32);
    return t >= 4294967296.0 ? 4294967296ull : (uint64_t)t;
  };
  const uint64_t ta = thr(a), tab = thr(a + b), tabc = thr(a + b + c);
  std::vector<uint64_t> out(samples);
  std::vector<uint8_t> keep(samples);
  parallel_for_dynamic(nblocks, 1, [&](uint64_t blk) {
    Xoshiro256ss rng(stream_seed(seed, blk));
    uint64_t beg = blk * kBlock, end = std::min(samples, beg + kBlock);
    for (uint64_t s = beg; s < end; ++s) {
      uint64_t u = 0, v = 0;
      uint64_t r = 0;
      for (int lvl = 0; lvl < scale; ++lvl) {
        if ((lvl & 1) == 0) r = rng.next();
        uint64_t x = (lvl & 1) ? (r >> 32) : (r & 0xFFFFFFFFull);
        uint64_t bit = 1ull << (scale - 1 - lvl);
        if (x < ta) {
        } else if (x < tab) {
          v |= bit;
        } else if (x < tabc) {
          u |= bit;
        } else {
          u |= bit;
          v |= bit;
        }
      }
      keep[s] = u != v;
      if (u > v) std::swap(u, v);
      out[s] = (u << 32) | v;
    }
  });
  keys.reserve(keys.size() + samples);
  for (uint64_t s = 0; s < samples; ++s)
    if (keep[s]) keys.push_back(out[s]);
}

namespace {
struct Alias {
  std::vector<double> prob;
  std::vector<uint32_t> alias;
  explicit Alias(const std::vector<double> &w) {
    const uint64_t n = w.size();
    prob.resize(n);
    alias.resize(n);
    double sum = 0;
    for (double x : w) sum += x;
    std::vector<double> scaled(n);
    std::vector<uint32_t> small, large;
    for (uint64_t i = 0; i < n; ++i) {
      scaled[i] = w[i] * (double)n / sum;
      (scaled[i] < 1.0 ? small : large).push_back((uint32_t)i);
    }
    while (!small.empty() && !large.empty()) {
      uint32_t s = small.back(), l = large.back();
      small.pop_back();
      prob[s] = scaled[s];
      alias[s] = l;
      scaled[l] = (scaled[l] + scaled[s]) - 1.0;
      if (scaled[l] < 1.0) {
        large.pop_back();
        small.push_back(l);
      }
    }
    for (uint32_t i : large) prob[i] = 1.0, alias[i] = i;
    for (uint32_t i : small) prob[i] = 1.0, alias[i] = i;
  }
  uint32_t sample(Xoshiro256ss &rng) const {
    uint64_t i = rng.next_below(prob.size());
    return rng.next_double() < prob[i] ? (uint32_t)i : alias[i];
  }
};
}  // namespace

void gen_chung_lu_planted(uint64_t n, double gamma, double mean_degree, uint32_t n_comm, uint32_t cmin,
                          uint32_t cmax, double p_in, uint64_t seed, std::vector<uint64_t> &keys) {
  const double beta = 1.0 / (gamma - 1.0);
  const double total = (double)n * mean_degree;
  const double cap = std::sqrt(total);
  std::vector<double> raw(n);
  for (uint64_t i = 0; i < n; ++i) raw[i] = std::pow((double)(i + 1), -beta);
  auto wsum = [&](double C) {
    double s = 0;
    for (uint64_t i = 0; i < n; ++i) s += std::min(C * raw[i], cap);
    return s;
  };
  double lo = 0, hi = 1;
  while (wsum(hi) < total && hi < 1e300) hi *= 2;
  for (int it = 0; it < 100; ++it) {
    double mid = 0.5 * (lo + hi);
    (wsum(mid) < total ? lo : hi) = mid;
  }
  std::vector<double> w(n);
  for (uint64_t i = 0; i < n; ++i) w[i] = std::min(hi * raw[i], cap);
  std::vector<double>().swap(raw);
  Alias alias(w);
  std::vector<double>().swap(w);

  const uint64_t samples = (uint64_t)std::llround(total / 2.0);
  const uint64_t kBlock = 1ull << 20;
  const uint64_t nblocks = (samples + kBlock - 1) / kBlock;
  std::vector<uint64_t> out(samples);
  std::vector<uint8_t> keep(samples);
  parallel_for_dynamic(nblocks, 1, [&](uint64_t blk) {
    Xoshiro256ss rng(stream_seed(seed, blk));
    uint64_t beg = blk * kBlock, end = std::min(samples, beg + kBlock);
    for (uint64_t s = beg; s < end; ++s) {
      uint64_t u = alias.sample(rng), v = alias.sample(rng);
      keep[s] = u != v;
      if (u > v) std::swap(u, v);
      out[s] = (u << 32) | v;
    }
  });
  keys.reserve(keys.size() + samples);
  for (uint64_t s = 0; s < samples; ++s)
    if (keep[s]) keys.push_back(out[s]);

  // Planted communities: stream (seed ^ tag, community index).
  std::vector<std::vector<uint64_t>> comm_edges(n_comm);
  parallel_for_dynamic(n_comm, 1, [&](uint64_t ci) {
    Xoshiro256ss rng(stream_seed(seed ^ 0xC0FFEEull, ci));
    uint32_t size = cmin + (uint32_t)rng.next_below((uint64_t)(cmax - cmin) + 1);
    if (size > n) size = (uint32_t)n;
    std::vector<uint32_t> mem;
    mem.reserve(size);
    while (mem.size() < size) {
      uint32_t x = (uint32_t)rng.next_below(n);
      if (std::find(mem.begin(), mem.end(), x) == mem.end()) mem.push_back(x);
    }
    auto &e = comm_edges[ci];
    for (uint32_t i = 0; i < size; ++i)
      for (uint32_t j = i + 1; j < size; ++j)
        if (rng.next_double() < p_in) {
          uint64_t a = mem[i], b = mem[j];
          if (a > b) std::swap(a, b);
          e.push_back((a << 32) | b);
        }
  });
  for (auto &e : comm_edges) keys.insert(keys.end(), e.begin(), e.end());
}

BatchGraphParams batch_graph_params(uint64_t seed, uint64_t index, uint32_t nmin, uint32_t nmax, double pmin,
                                    double pmax) {
  Xoshiro256ss rng(stream_seed(seed, index));
  BatchGraphParams bp;
  bp.n = nmin + (uint32_t)rng.next_below((uint64_t)(nmax - nmin) + 1);
  bp.p = pmin + (pmax - pmin) * rng.next_double();
  bp.seed = rng.next();
  return bp;
}

}  // namespace rdmce
