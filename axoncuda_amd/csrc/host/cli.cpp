// cli.cpp - `hgraph_gpu`: the drop-in command line of the reference partitioner
// (sources/main.cu:38-940, sources/runconfig.cpp).  Same options, defaults, printed
// labels and exit codes; the device section runs through the libhgp C ABI.
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>

#include "hgp.h"
#include "model.hpp"

namespace {

void print_help() {
  std::printf(
      "Usage:\n"
      "  prog -r <input_file> [-c <constr>] [-s <output_file>] [-p <part_file>]\n"
      "  prog -r <input_file> [-k <k> <\xce\xb5>] [-s <output_file>] [-p <part_file>]\n"
      "  prog -h\n\n"
      "Options:\n"
      "  -r <file>   Reload hypergraph from file\n"
      "  -s <file>   Save partitioned hypergraph to file\n"
      "  -p <file>   Save the partitioning to file (one line per node, containing its partition id)\n"
      "  -c <name>   Preconfigured constraints set to use (valid ones: truenorth, loihi64, loihi84, loihi1024 - default is loihi64)\n"
      "  -m <> <> <> Constraints set to use, in order: max part. size, max part. distinct inbound hedges, max num. of part.s (overrides '-c')\n"
      "  -k <k> <\xce\xb5>  K-way balanced constraints set to use (overrides '-c' and '-m')\n"
      "  -om <mult>  Set the deduplication oversized segment size multiplier (increase to avoid the 'GM hash-set full!' assert)\n"
      "  -cnc <num>  Set the count of candidates proposed per node during coarsening\n"
      "  -rfr <num>  Set the number of refinement repetitions per level\n"
      "  -smh <lvl>  Recover device memory by migrating to the host the pre-coarsening hypergraph up the provided level\n"
      "  -dtc        When set, construct touching sets on the device, rather than on the host (faster - uses more device memory)\n"
      "  -ipm        When set, initial partitions are greedily merged to reduce their number (and save memory - minor quality loss)\n"
      "  -ip <where> k-way initial partitioner without mtkahypar: 'auto' (default: 'device' on levels of 2^21 pins or more), 'device' (multistart label propagation) or 'host' (recursive bisection)\n"
      "  -lp <num>   k-way: label-propagation rounds before each level's FM rounds (default 0)\n"
      "  -tune <file> Launch parameters: read them from the file, or autotune this device and write them there\n"
      "  -v <lvl>    Set the verbosity level: 0 metrics only, 1 steps and phases, 2 kernel launches, 3  \n"
      "  -h          Show this help message\n");
}

struct Named {
  const char* key;
  const char* name;
  uint32_t omega, delta, maxp;
};
// Constraints presets (sources/constr.cpp:195-229)
const Named kPresets[] = {{"loihi64", "Loihi Large", 1024, 4096, 4096},
                          {"loihi84", "Loihi Jin 84", 4096, 1024 * 64, 7056},
                          {"loihi1024", "Loihi Jin 1024", 4096, 1024 * 64, 1048576},
                          {"truenorth", "TrueNorth", 256, 256, 4096}};

unsigned long stoul_or_die(const char* s, const char* what) {
  char* end = nullptr;
  errno = 0;
  unsigned long v = std::strtoul(s, &end, 10);
  if (end == s || errno) {
    std::fprintf(stderr, "Error: invalid value for %s: %s\n", what, s);
    std::exit(1);
  }
  return v;
}

float stof_or_die(const char* s, const char* what) {
  char* end = nullptr;
  float v = std::strtof(s, &end);
  if (end == s) {
    std::fprintf(stderr, "Error: invalid value for %s: %s\n", what, s);
    std::exit(1);
  }
  return v;
}

double file_mb(const std::string& p) {
  struct stat st;
  if (::stat(p.c_str(), &st) != 0) return 0.0;
  return (double)st.st_size / (1 << 20);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 1) {
    print_help();
    return 0;
  }
  enum { C_NONE, C_NAME, C_MANL, C_KWAY } ctype = C_NONE;
  std::string load_path, save_path, part_path, cname, tune_path;
  hgp_params prm;
  hgp_default_params(&prm);
  uint32_t m_omega = 0, m_delta = 0, m_maxp = 0;
  bool verbose_info = true, verbose_err = true;
  int verbosity = 1;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-h") {
      print_help();
      return 0;
    } else if (a == "-r") {
      if (i + 1 >= argc) { std::fprintf(stderr, "Error: -r requires a file path\n"); return 1; }
      load_path = argv[++i];
    } else if (a == "-s") {
      if (i + 1 >= argc) { std::fprintf(stderr, "Error: -s requires a file path\n"); return 1; }
      save_path = argv[++i];
    } else if (a == "-p") {
      if (i + 1 >= argc) { std::fprintf(stderr, "Error: -p requires a file path\n"); return 1; }
      part_path = argv[++i];
    } else if (a == "-c") {
      if (i + 1 >= argc) { std::fprintf(stderr, "Error: -c requires a config name\n"); return 1; }
      ctype = C_NAME;
      cname = argv[++i];
      prm.mode = HGP_MODE_INCC;
    } else if (a == "-m") {
      if (i + 3 >= argc) { std::fprintf(stderr, "Error: -m requires integer values for the three constraints\n"); return 1; }
      ctype = C_MANL;
      m_omega = (uint32_t)stoul_or_die(argv[++i], "-m");
      m_delta = (uint32_t)stoul_or_die(argv[++i], "-m");
      m_maxp = (uint32_t)stoul_or_die(argv[++i], "-m");
      prm.mode = HGP_MODE_INCC;
    } else if (a == "-k") {
      if (i + 2 >= argc) { std::fprintf(stderr, "Error: -k requires values for 'k' and '\xce\xb5'\n"); return 1; }
      ctype = C_KWAY;
      prm.kway = (uint32_t)stoul_or_die(argv[++i], "-k");
      prm.epsilon = stof_or_die(argv[++i], "-k");
      prm.mode = HGP_MODE_KWAY;
    } else if (a == "-om") {
      if (i + 1 >= argc) { std::fprintf(stderr, "Error: -om requires a float value\n"); return 1; }
      prm.oversized_multiplier = stof_or_die(argv[++i], "-om");
    } else if (a == "-cnc") {
      if (i + 1 >= argc) { std::fprintf(stderr, "Error: -cnc requires a positive integer value\n"); return 1; }
      prm.candidates_count = (uint32_t)stoul_or_die(argv[++i], "-cnc");
      if (prm.candidates_count > 16) { std::fprintf(stderr, "Error: -cnc must be less or equal to 16\n"); return 1; }
    } else if (a == "-rfr") {
      if (i + 1 >= argc) { std::fprintf(stderr, "Error: -rfr requires a positive integer value\n"); return 1; }
      prm.refine_repeats = (uint32_t)stoul_or_die(argv[++i], "-rfr");
    } else if (a == "-smh") {
      if (i + 1 >= argc) { std::fprintf(stderr, "Error: -smh requires a positive integer value\n"); return 1; }
      prm.save_memory_up_to_level = (uint32_t)stoul_or_die(argv[++i], "-smh");
    } else if (a == "-dtc") {
      prm.device_touching_construction = 1;
    } else if (a == "-ipm") {
      prm.initial_partitions_merge = 1;
    } else if (a == "-ip") {
      if (i + 1 >= argc) { std::fprintf(stderr, "Error: -ip requires 'auto', 'host' or 'device'\n"); return 1; }
      const std::string w = argv[++i];
      if (w == "host") prm.initial_partitioner = HGP_IP_HOST;
      else if (w == "device") prm.initial_partitioner = HGP_IP_DEVICE;
      else if (w == "auto") prm.initial_partitioner = HGP_IP_AUTO;
      else { std::fprintf(stderr, "Error: -ip requires 'auto', 'host' or 'device'\n"); return 1; }
    } else if (a == "-tune") {
      if (i + 1 >= argc) { std::fprintf(stderr, "Error: -tune requires a file path\n"); return 1; }
      tune_path = argv[++i];
    } else if (a == "-lp") {
      if (i + 1 >= argc) { std::fprintf(stderr, "Error: -lp requires a positive integer value\n"); return 1; }
      prm.lp_rounds = (uint32_t)stoul_or_die(argv[++i], "-lp");
    } else if (a == "-v") {
      if (i + 1 >= argc) { std::fprintf(stderr, "Error: -v requires a positive value between 0 and 3\n"); return 1; }
      verbosity = (int)stoul_or_die(argv[++i], "-v");
      if (verbosity < 0 || verbosity > 3) { std::fprintf(stderr, "Error: -v must be between 0 and 3 (extremes included) \n"); return 1; }
      verbose_info = verbosity > 0;
      verbose_err = verbosity > 0;
    } else {
      std::fprintf(stderr, "Unknown option: %s\n", a.c_str());
      return 1;
    }
  }
  prm.verbose = verbosity > 1 ? verbosity - 1 : 0;
  if (prm.candidates_count == 0) { std::fprintf(stderr, "Error: -cnc requires a positive integer value\n"); return 1; }

  // The device context (the HIP runtime's start-up) is created on a thread of its own while
  // the file is read and deduplicated; the device is first used after that.
  struct CtxStart {
    hgp_ctx* ctx = nullptr;
    hgp_status st = HGP_E_NODEVICE;
    std::thread th;
    CtxStart() : th([this] { st = hgp_ctx_create(0, 0, &ctx); }) {}
    hgp_status take(hgp_ctx** out) {
      if (th.joinable()) th.join();
      *out = ctx;
      ctx = nullptr;
      return st;
    }
    ~CtxStart() {
      if (th.joinable()) th.join();
      if (ctx) hgp_ctx_destroy(ctx);
    }
  } ctx_start;

  // load (sources/runconfig.cpp:153-184)
  hgp_hgraph* hg = nullptr;
  if (!load_path.empty()) {
    struct stat st;
    if (::stat(load_path.c_str(), &st) != 0 || !S_ISREG(st.st_mode)) {
      std::fprintf(stderr, "Error loading file: Failed to load hypergraph, the provided path is not a file.\n");
      return 1;
    }
    const char* fmt = load_path.size() > 4 && load_path.compare(load_path.size() - 4, 4, ".hgr") == 0 ? "hMETIS"
                      : load_path.size() > 4 && load_path.compare(load_path.size() - 4, 4, ".snn") == 0 ? "SNN"
                      : load_path.size() > 4 && load_path.compare(load_path.size() - 4, 4, ".axh") == 0 ? "AXH"
                                                                                                         : nullptr;
    if (!fmt) {
      std::fprintf(stderr, "Error loading file: Failed to load hypergraph, unsupported file format (supported: '.hgr', '.snn', '.axh').\n");
      return 1;
    }
    std::printf("Loading hypergraph from: %s (%s format) ...\n", load_path.c_str(), fmt);
    std::printf("Hypergraph file size: %.1f MB\n", file_mb(load_path));
    if (hgp_hgraph_load(load_path.c_str(), 0, &hg) != HGP_OK) return 1;
  } else {
    std::fprintf(stderr, "WARNING, no hypergraph provided (-r), performing a dry-run !!\n");
    if (hgp_hgraph_from_csr(0, 0, (const uint64_t[]){0}, nullptr, nullptr, nullptr, 0, &hg) != HGP_OK) return 1;
  }
  uint32_t n = 0, e = 0;
  uint64_t p = 0;
  hgp_hgraph_sizes(hg, &n, &e, &p);

  // constraints (sources/runconfig.cpp:186-217)
  std::string cfg_name;
  if (ctype == C_KWAY) {
    char eps[32];
    std::snprintf(eps, sizeof eps, "%.3f", prm.epsilon);
    cfg_name = std::to_string(prm.kway) + "-way " + eps + " balanced";
    prm.max_nodes_per_part = (uint32_t)std::ceil((1 + prm.epsilon) * (float)n / prm.kway);
    prm.max_inbound_per_part = 2147483647u;
    prm.max_parts = prm.kway;
  } else if (ctype == C_MANL) {
    if (m_omega == 0) { std::fprintf(stderr, "Error: the 1st constraint (max partition size) must be a positive integer \n"); return 1; }
    if (m_delta == 0) { std::fprintf(stderr, "Error: the 2nd constraint (max distinct inbound hedge per partition) must be a positive integer \n"); return 1; }
    if (m_maxp == 0) { std::fprintf(stderr, "Error: the 3rd constraint (max number of partitions) must be a positive integer \n"); return 1; }
    cfg_name = "manual";
    prm.max_nodes_per_part = m_omega;
    prm.max_inbound_per_part = m_delta;
    prm.max_parts = m_maxp;
  } else {
    const Named* sel = &kPresets[0];
    if (ctype == C_NAME) {
      bool found = false;
      for (const Named& x : kPresets)
        if (cname == x.key) {
          sel = &x;
          found = true;
        }
      if (!found) std::fprintf(stderr, "WARNING, constraints name (-c %s) not recognized, using loihi64 !!\n", cname.c_str());
    } else {
      std::fprintf(stderr, "WARNING, no constraints provided (-c, -m, -k), using loihi64 !!\n");
    }
    cfg_name = sel->name;
    prm.max_nodes_per_part = sel->omega;
    prm.max_inbound_per_part = sel->delta;
    prm.max_parts = sel->maxp;
  }

  // statistics (sources/main.cu:54-70)
  std::vector<uint64_t> off((size_t)e + 1);
  std::vector<float> w(e);
  hgp_hgraph_copy_out(hg, off.data(), nullptr, nullptr, w.data());
  double total_w = 0;
  for (uint32_t i = 0; i < e; ++i) total_w += (double)w[i] * (double)(off[i + 1] - off[i] - 1);
  std::printf("Loaded hypergraph:\n");
  std::printf("  Nodes:      %u\n", n);
  std::printf("  Hyperedges: %u\n", e);
  std::printf("  Total pins: %llu\n", (unsigned long long)p);
  std::printf("  Total connections weight: %.3f\n", total_w);
  std::printf("Using constraints \"%s\":\n", cfg_name.c_str());
  std::printf("  Nodes per partition:         %u\n", prm.max_nodes_per_part);
  std::printf("  Inbound hedge per partition: %u\n", prm.max_inbound_per_part);
  std::printf("  Maximum partitions:          %u\n", prm.max_parts);
  std::printf("Using settings:\n");
  std::printf("  Candidates count:            %u\n", prm.candidates_count);
  std::printf("  Refinement repetitions:      %u\n", prm.refine_repeats);
  std::printf("  Oversized multiplier factor: %.2f\n", prm.oversized_multiplier);
  std::printf("  Host memory spill levels:    %u\n", prm.save_memory_up_to_level);
  std::printf("  Flags: %s%s\n", prm.device_touching_construction ? "dtc " : "", prm.initial_partitions_merge ? "ipm " : "");

  hgp_ctx* ctx = nullptr;
  hgp_status s = ctx_start.take(&ctx);
  if (s != HGP_OK) {
    std::fprintf(stderr, "ERROR: no usable gfx950 device (status %d)\n", s);
    return s;
  }
  std::printf("HIP device:\n  Using device 0 (gfx950)\n");
  if (!tune_path.empty()) {  // launch parameters only: the partition does not depend on them
    hgp_tuning t;
    if (hgp_tuning_load(tune_path.c_str(), &t) == HGP_OK) {
      hgp_tuning_set(ctx, &t);
      std::printf("Launch parameters read from %s\n", tune_path.c_str());
    } else if (hgp_autotune(ctx, 0, &t) == HGP_OK && hgp_tuning_save(&t, tune_path.c_str()) == HGP_OK) {
      std::printf("Launch parameters autotuned and written to %s\n", tune_path.c_str());
    } else {
      std::fprintf(stderr, "WARNING: could not autotune or write %s, using the built-in launch parameters\n",
                   tune_path.c_str());
    }
  }

  // rule-2 dedup before the device section (sources/main.cu:125)
  hgp_hgraph* dd = nullptr;
  {
    std::vector<uint32_t> pins(p), srcc(e);
    hgp_hgraph_copy_out(hg, off.data(), pins.data(), srcc.data(), w.data());
    if (hgp_hgraph_from_csr(n, e, off.data(), pins.data(), srcc.data(), w.data(), 1, &dd) != HGP_OK) return 1;
  }
  auto t0 = std::chrono::high_resolution_clock::now();
  std::vector<uint32_t> part(n);
  hgp_run_stats st{};
  if (n > 0) {
    s = hgp_partition(ctx, &prm, dd, part.data(), &st);
    if (s != HGP_OK) {
      std::fprintf(stderr, "ERROR: partitioning failed (status %d): %s\n", s, hgp_ctx_last_error(ctx));
      hgp_ctx_destroy(ctx);
      return s;
    }
  }
  auto t1 = std::chrono::high_resolution_clock::now();
  const double host_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  std::printf("Total device core execution time: %.3f ms\n", st.core_ms);
  std::printf("Total device execution time: %.3f ms\n", st.device_ms);
  std::printf("Total host execution time: %.3f ms\n", host_ms);

  int32_t valid = 0;
  hgp_check_validity(dd, part.data(), prm.max_nodes_per_part, prm.max_inbound_per_part, prm.max_parts, &valid);
  int rc = 0;
  if (valid == 1) {
    double km1 = 0, cut = 0, soed = 0;
    hgp_hgraph_metrics(dd, part.data(), &km1, &cut, &soed);
    std::printf("Partitioned hypergraph:\n");
    std::printf("  Nodes:         %u\n", st.num_partitions);
    std::printf("  Cut-net:       %.3f\n", cut);
    std::printf("  Connectivity:  %.3f\n", km1);
    std::printf("  SOED:          %.3f\n", soed);
    if (!save_path.empty()) {
      std::printf("Saving partitioned hypergraph to: %s ...\n", save_path.c_str());
      if (hgp_save_partitioned(dd, part.data(), save_path.c_str()) != HGP_OK) {
        std::fprintf(stderr, "Error saving file: %s\n", save_path.c_str());
        rc = 1;
      } else {
        std::printf("Partitioned hypergraph saved to %s\n", save_path.c_str());
      }
    }
    if (!part_path.empty()) {
      std::printf("Saving partitioning to: %s (each node's partition id on its line by node idx) ...\n", part_path.c_str());
      if (hgp_save_partitioning(part_path.c_str(), part.data(), n) != HGP_OK) {
        std::fprintf(stderr, "Error saving file: %s\n", part_path.c_str());
        rc = 1;
      } else {
        std::printf("Partitioning saved to %s\n", part_path.c_str());
      }
    }
  } else {
    std::fprintf(stderr, "ERROR, invalid partitining !!\n");
    rc = 1;
  }
  (void)verbose_info;
  (void)verbose_err;
  hgp_hgraph_free(dd);
  hgp_hgraph_free(hg);
  hgp_ctx_destroy(ctx);
  return rc;
}
