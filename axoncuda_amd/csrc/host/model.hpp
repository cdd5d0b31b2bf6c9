// model.hpp - host hypergraph model (reference includes/hgraph.hpp HyperGraph):
// CSR of pins with sources first, per-hedge source count and float weight.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

struct hgp_hgraph {
  uint32_t n = 0;
  std::vector<uint64_t> off{0};  // hedge offsets (E+1)
  std::vector<uint32_t> pins;
  std::vector<uint32_t> srcc;    // sources per hedge (stored first)
  std::vector<float> w;
  std::vector<uint32_t> vweights;  // .hgr vertex weights when present (not used by the pipeline)
  uint32_t e() const { return (uint32_t)srcc.size(); }
};

namespace hgp {

// Returns an empty string on success, the error text otherwise.
std::string load_hgraph(const std::string& path, hgp_hgraph& hg);
void dedup_rule2(hgp_hgraph& hg);
std::string save_hgraph(const hgp_hgraph& hg, const std::string& path);

// quotient hypergraph getPartitionsHypergraph(part, 2, true) (includes/hgraph.hpp:286-376)
hgp_hgraph partitions_hypergraph(const hgp_hgraph& hg, const std::vector<uint32_t>& part);

struct Metrics {
  double connectivity = 0, cutnet = 0, soed = 0;
  uint32_t quotient_hedges = 0;
  uint64_t quotient_pins = 0;
};
Metrics metrics(const hgp_hgraph& hg, const std::vector<uint32_t>& part);
// 1 valid, 0 invalid, -1 partition ids not 0..count-1; `why` receives the reason
int validity(const hgp_hgraph& hg, const std::vector<uint32_t>& part, uint32_t omega, uint32_t delta, uint32_t max_parts,
             std::string* why);

void gen_random(uint32_t n, uint32_t e, uint32_t max_size, uint64_t seed, hgp_hgraph& hg);
void gen_netlist(uint32_t n, uint32_t e, double alpha, uint32_t max_size, uint32_t radius, double frac_local,
                 uint64_t seed, hgp_hgraph& hg);
void gen_snn(uint32_t layers, uint32_t per_layer, uint32_t grid, uint32_t fanout, double sigma, uint64_t seed,
             hgp_hgraph& hg);

}  // namespace hgp
