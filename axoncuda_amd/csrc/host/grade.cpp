// grade.cpp - `hgrade`: partition evaluator with the reference grader's command line and
// printed metrics (grading/grade.cpp:158-360, partition half only): loads the hypergraph
// and a partition file, checks validity (sources/constr.cpp:90-147) and prints cut-net,
// Connectivity (lambda-1) and SOED over getPartitionsHypergraph(part, 2, true).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "model.hpp"

namespace {

bool load_partitioning(const std::string& path, uint32_t n, std::vector<uint32_t>& part, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "Cannot open partitioning file.";
    return false;
  }
  part.assign(n, 0);
  std::string line;
  uint32_t idx = 0;
  while (std::getline(f, line)) {
    if (idx >= n) {
      err = "Partitioning file has more entries than the hypergraph has nodes.";
      return false;
    }
    std::istringstream ls(line);
    uint64_t p = 0;
    if (!(ls >> p)) {
      err = "Invalid partition id at node " + std::to_string(idx) + ".";
      return false;
    }
    ls >> std::ws;
    if (!ls.eof()) {
      err = "Unexpected trailing data at node " + std::to_string(idx) + ".";
      return false;
    }
    if (p > 0xFFFFFFFFull) {
      err = "Partition id at node " + std::to_string(idx) + " exceeds uint32_t range.";
      return false;
    }
    part[idx++] = (uint32_t)p;
  }
  if (idx != n) {
    err = "Partitioning file has " + std::to_string(idx) + " entries, but the hypergraph has " + std::to_string(n) + " nodes.";
    return false;
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 1) {
    std::printf("Usage:\n  prog -r <hgraph_file> -prt <partitioning_file> (-k-prt <k> <eps> | -m-prt <> <> <> | -c-prt <name>) [-s <file>]\n");
    return 0;
  }
  std::string load_path, part_path, save_path, cname;
  int ctype = 0;  // 0 none, 1 name, 2 manual, 3 kway
  uint32_t omega = 0, delta = 0, maxp = 0, k = 0;
  float eps = 0;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto need = [&](int cnt) {
      if (i + cnt >= argc) {
        std::fprintf(stderr, "Error: %s requires %d value(s)\n", a.c_str(), cnt);
        std::exit(1);
      }
    };
    if (a == "-r") { need(1); load_path = argv[++i]; }
    else if (a == "-prt") { need(1); part_path = argv[++i]; }
    else if (a == "-s") { need(1); save_path = argv[++i]; }
    else if (a == "-c-prt") { need(1); ctype = 1; cname = argv[++i]; }
    else if (a == "-m-prt") { need(3); ctype = 2; omega = std::stoul(argv[++i]); delta = std::stoul(argv[++i]); maxp = std::stoul(argv[++i]); }
    else if (a == "-k-prt") { need(2); ctype = 3; k = std::stoul(argv[++i]); eps = std::stof(argv[++i]); }
    else if (a == "-h") { std::printf("Usage: prog -r <hgraph_file> -prt <partitioning_file> ...\n"); return 0; }
    else { std::fprintf(stderr, "Unknown option: %s\n", a.c_str()); return 1; }
  }
  if (ctype == 0) {
    std::printf("Evaluation task:\n  -> no task provided (no option provided among -c-prt, -m-prt, -k-prt)\n");
    return 0;
  }
  if (load_path.empty()) { std::fprintf(stderr, "Error: -r is required to load the hypergraph to evaluate.\n"); return 1; }
  if (part_path.empty()) { std::fprintf(stderr, "Error: partitioning evaluation requires -prt <partitioning_file>.\n"); return 1; }
  hgp_hgraph hg;
  std::string err = hgp::load_hgraph(load_path, hg);
  if (!err.empty()) { std::fprintf(stderr, "Error loading file: %s\n", err.c_str()); return 1; }
  if (ctype == 3) {
    omega = (uint32_t)std::ceil((1 + eps) * (float)hg.n / k);
    delta = 2147483647u;
    maxp = k;
  } else if (ctype == 1) {
    struct P { const char* n; uint32_t o, d, m; } presets[] = {
        {"loihi64", 1024, 4096, 4096}, {"loihi84", 4096, 65536, 7056}, {"loihi1024", 4096, 65536, 1048576}, {"truenorth", 256, 256, 4096}};
    bool found = false;
    for (auto& x : presets)
      if (cname == x.n) { omega = x.o; delta = x.d; maxp = x.m; found = true; }
    if (!found) { std::fprintf(stderr, "WARNING, constraints name (-c %s) not recognized, aborting !!\n", cname.c_str()); return 1; }
  } else if (omega == 0 || delta == 0 || maxp == 0) {
    std::fprintf(stderr, "Error: constraints must be positive integers\n");
    return 1;
  }
  std::printf("Loaded hypergraph:\n  Nodes:      %u\n  Hyperedges: %u\n  Total pins: %llu\n", hg.n, hg.e(),
              (unsigned long long)hg.pins.size());
  std::vector<uint32_t> part;
  if (!load_partitioning(part_path, hg.n, part, err)) { std::fprintf(stderr, "Error loading partitioning: %s\n", err.c_str()); return 1; }
  std::string why;
  int v = hgp::validity(hg, part, omega, delta, maxp, &why);
  if (v != 1) {
    std::printf("INVALID PARTITIONING: %s\n", why.c_str());
    std::fprintf(stderr, "ERROR, invalid partitioning !!\n");
    return 1;
  }
  hgp::Metrics m = hgp::metrics(hg, part);
  uint32_t np = 0;
  for (uint32_t x : part) np = std::max(np, x + 1);
  std::printf("Partitioned hypergraph metrics:\n");
  std::printf("  Nodes:         %u\n", np);
  std::printf("  Total pins:    %llu\n", (unsigned long long)m.quotient_pins);
  std::printf("  Cut-net:       %.3f\n", m.cutnet);
  std::printf("  Connectivity:  %.3f\n", m.connectivity);
  std::printf("  SOED:          %.3f\n", m.soed);
  if (!save_path.empty()) {
    err = hgp::save_hgraph(hgp::partitions_hypergraph(hg, part), save_path);
    if (!err.empty()) { std::fprintf(stderr, "Error saving file: %s\n", err.c_str()); return 1; }
  }
  return 0;
}
