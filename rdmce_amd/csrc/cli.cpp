// rdmce: command-line driver with the reference's flags and stdout lines
// (RDMCE/main.cc:13-199, RDMCE/src/mce_gpu.cu:14-22), so its experiment
// scripts (scripts/analyze.sh:97 reads field 4 of "MCE GPU Time:") can
// drive it unchanged.
//
//   rdmce -i <graph.txt|graph.bin> [-o a|d|deg|unchange] [-d id] [-c]
//         [--vertex] [--tau N] [--waves N] [--hist] [--hash] [--list FILE]
//
// -d takes one device id (this build runs on one GPU; a list is rejected).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rdmce.h"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

[[noreturn]] void usage(const char *prog) {
  std::fprintf(stderr,
               "\033[33mUsage: %s -i <input_graph_file> -o <order_option> -d <device_id> [-c]\n"
               "Order options: a (ascending), d (descending), deg (degeneracy),  unchange (unchanged)\n"
               "Extra: --vertex (vertex roots) --tau N (sharing bound) --waves N (waves per CU)\n"
               "       --hist --hash --list FILE\n"
               "\033[36me.g., %s -i RDMCE/data/zachary.txt -o deg -d 0\033[0m\n",
               prog, prog);
  std::exit(1);
}

[[noreturn]] void die(const char *what) {
  std::fprintf(stderr, "%s: %s\n", what, mce_last_error());
  std::exit(1);
}

bool ends_with(const std::string &s, const char *suf) {
  size_t n = std::strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

}  // namespace

int main(int argc, char **argv) {
  std::string input, list_path;
  int order = MCE_ORDER_DEG, device = 0, tau = 0, waves = 0, root_mode = MCE_ROOT_EDGE;
  bool convert_only = false, want_hist = false, want_hash = false;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) usage(argv[0]);
      return argv[++i];
    };
    if (a == "-i") input = next();
    else if (a == "-o") {
      std::string o = next();
      if (o == "a") order = MCE_ORDER_ASC;
      else if (o == "d") order = MCE_ORDER_DES;
      else if (o == "deg") order = MCE_ORDER_DEG;
      else if (o == "unchange") order = MCE_ORDER_UNCHANGED;
      else usage(argv[0]);
    } else if (a == "-d") {
      std::string ids = next();
      if (ids.find(',') != std::string::npos) {
        std::fprintf(stderr, "this build runs on one GPU: give a single device id\n");
        return 1;
      }
      char *end = nullptr;
      long v = std::strtol(ids.c_str(), &end, 10);
      if (!*ids.c_str() || *end || v < 0) {
        std::fprintf(stderr, "invalid device ID: %s\n", ids.c_str());
        return 1;
      }
      device = (int)v;
    } else if (a == "-c") convert_only = true;
    else if (a == "--vertex") root_mode = MCE_ROOT_VERTEX;
    else if (a == "--tau") tau = std::atoi(next().c_str());
    else if (a == "--waves") waves = std::atoi(next().c_str());
    else if (a == "--hist") want_hist = true;
    else if (a == "--hash") want_hash = true;
    else if (a == "--list") list_path = next();
    else usage(argv[0]);
  }
  if (input.empty()) usage(argv[0]);

  static const char *order_names[] = {"Unchanged", "Ascending", "Descending", "Degeneracy"};
  std::printf("==================== Parameters ====================\n");
  std::printf("Device ID: %d\n", device);
  std::printf("Input file: %s\n", input.c_str());
  std::printf("Order type: %s\n", order_names[order]);
  std::printf("Levels: %s\n", root_mode == MCE_ROOT_EDGE ? "L2" : "L1");
  std::printf("====================================================\n");

  mce_graph *g = nullptr;
  double t0 = now_s();
  const bool is_bin = ends_with(input, ".bin");
  if (is_bin) {
    if (mce_graph_load_bin(input.c_str(), &g)) die("load failed");
  } else {
    if (mce_graph_load_coo_text(input.c_str(), &g)) die("load failed");
  }
  std::printf("Load time: %gs\n", now_s() - t0);
  if (!is_bin) {
    t0 = now_s();
    if (mce_graph_order(g, order)) die("ordering failed");
    std::printf("Sort time: %gs\n", now_s() - t0);
    if (convert_only) {
      char out[4096];
      if (mce_graph_bin_path(input.c_str(), out, sizeof(out)) || mce_graph_store_bin(g, out)) die("store failed");
      std::printf("preprocess %s done\n", input.c_str());
    }
  }
  mce_graph_info info;
  mce_graph_get_info(g, &info);
  std::printf("vertices: %llu \tedges: %llu\n", (unsigned long long)info.n, (unsigned long long)info.m);
  std::printf("max degree: %llu \tdegeneracy: %llu\n", (unsigned long long)info.max_degree,
              (unsigned long long)info.degeneracy);
  if (convert_only) {
    std::printf("Convert only mode (-c) enabled, exiting after conversion.\n");
    mce_graph_free(g);
    return 0;
  }

  mce_options opt;
  mce_options_init(&opt);
  opt.device = device;
  opt.root_mode = root_mode;
  opt.share_bound = tau;
  opt.waves_per_cu = waves;
  opt.want_list = list_path.empty() ? 0 : 1;
  if (opt.want_list) opt.list_capacity = 1ull << 28;
  mce_result *r = nullptr;
  t0 = now_s();
  if (mce_solve(g, &opt, &r)) die("enumeration failed");
  double elapsed = now_s() - t0;
  std::printf("\nMC numbers: %llu\n", (unsigned long long)r->count);
  std::printf("MCE GPU Time: %g\n", elapsed);
  if (want_hash) std::printf("MC hash: %llu\n", (unsigned long long)r->hash);
  if (want_hist) {
    std::printf("MC histogram:");
    for (uint32_t k = 0; k < r->hist_len; ++k)
      if (r->hist[k]) std::printf(" %u:%llu", k, (unsigned long long)r->hist[k]);
    std::printf("\n");
  }
  if (!list_path.empty()) {
    FILE *f = std::fopen(list_path.c_str(), "w");
    if (!f) die("cannot open clique list file");
    for (uint64_t i = 0; i < r->list_len;) {
      uint32_t k = r->list[i];
      for (uint32_t j = 0; j < k; ++j) std::fprintf(f, j ? " %u" : "%u", r->list[i + 1 + j]);
      std::fprintf(f, "\n");
      i += k + 1;
    }
    std::fclose(f);
  }
  mce_result_free(r);
  mce_graph_free(g);
  return 0;
}
