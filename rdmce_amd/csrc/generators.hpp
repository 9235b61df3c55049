// Deterministic synthetic graph generators.  Every generator emits raw
// canonical pairs (min << 32 | max, ids as generated, self loops removed);
// graph_from_keys() then deduplicates and relabels them exactly as the COO
// loader does.  Output depends only on the parameters and the seed, never on
// the host thread count (work is cut into fixed blocks with their own
// xoshiro256** streams).
#pragma once
#include <cstdint>
#include <vector>

namespace rdmce {

// Erdos-Renyi G(n, p): every pair u < v kept with probability p, pairs
// visited in (u, v) lexicographic order from one xoshiro256** stream.
void gen_er(uint32_t n, double p, uint64_t seed, std::vector<uint64_t> &keys);

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004): edge_factor * 2^scale samples;
// each sample descends `scale` levels choosing quadrant (0,0),(0,1),(1,0),(1,1)
// with probabilities a, b, c, 1-a-b-c (32-bit thresholds).  Samples are cut
// into blocks of 2^20 with one stream per block.
void gen_rmat(int scale, uint64_t edge_factor, double a, double b, double c, uint64_t seed,
              std::vector<uint64_t> &keys);

// Chung-Lu power-law graph plus planted dense communities.
//   weights  w_i = min(C * (i+1)^(-1/(gamma-1)), K), K = sqrt(n * mean_degree),
//            C found by bisection so that sum w = n * mean_degree;
//   edges    round(n * mean_degree / 2) samples, both endpoints drawn
//            independently with probability w_i / sum w (Vose alias table);
//   communities: n_comm groups, size uniform in [cmin, cmax], members drawn
//            uniformly without replacement from [0, n), each member pair
//            joined with probability p_in.
void gen_chung_lu_planted(uint64_t n, double gamma, double mean_degree, uint32_t n_comm, uint32_t cmin,
                          uint32_t cmax, double p_in, uint64_t seed, std::vector<uint64_t> &keys);

// Parameters of graph i of a batch of independent G(n, p): n uniform in
// [nmin, nmax], p uniform in [pmin, pmax), edge seed derived from (seed, i).
struct BatchGraphParams {
  uint32_t n;
  double p;
  uint64_t seed;
};
BatchGraphParams batch_graph_params(uint64_t seed, uint64_t index, uint32_t nmin, uint32_t nmax, double pmin,
                                    double pmax);

}  // namespace rdmce
