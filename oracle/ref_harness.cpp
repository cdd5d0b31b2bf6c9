// ref_harness.cpp - extern "C" driver over the reference's own host C++ (test infrastructure).
//
// Compiled by oracle/Makefile directly against the reference sources where they lie
// (includes/hgraph.hpp, includes/constr.hpp, sources/constr.cpp), into oracle/_ref/ only.
// It adds no code of the reference's and replaces none of its headers: it only calls the
// reference loaders, rule-2 dedup, quotient-hypergraph metrics and validity check so that
// tests can pin the oracle's restatement of those functions against the real thing.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "constr.hpp"
#include "hgraph.hpp"

using hgraph::HyperGraph;

extern "C" {

void* ref_load(const char* path, int dedup) {
  try {
    std::string p(path);
    HyperGraph* hg = nullptr;
    if (p.size() > 4 && p.substr(p.size() - 4) == ".hgr") hg = new HyperGraph(HyperGraph::loadhMETIS(p, false));
    else if (p.size() > 4 && p.substr(p.size() - 4) == ".snn") hg = new HyperGraph(HyperGraph::loadSNN(p, false));
    else if (p.size() > 4 && p.substr(p.size() - 4) == ".axh") hg = new HyperGraph(HyperGraph::loadAXH(p, false));
    else return nullptr;
    if (dedup) hg->deduplicateHyperedges(2, false);
    return hg;
  } catch (...) {
    return nullptr;
  }
}

void ref_sizes(void* h, uint32_t* n, uint32_t* e, uint64_t* p) {
  HyperGraph* hg = (HyperGraph*)h;
  *n = hg->nodes();
  *e = (uint32_t)hg->hedges().size();
  *p = hg->hedgesFlat().size();
}

void ref_copy(void* h, uint64_t* off, uint32_t* pins, uint32_t* srcc, float* w) {
  HyperGraph* hg = (HyperGraph*)h;
  const auto& he = hg->hedges();
  for (size_t i = 0; i < he.size(); ++i) {
    off[i] = he[i].offset();
    srcc[i] = he[i].src_count();
    w[i] = he[i].weight();
  }
  off[he.size()] = hg->hedgesFlat().size();
  const auto& f = hg->hedgesFlat();
  for (size_t i = 0; i < f.size(); ++i) pins[i] = f[i];
}

void ref_free(void* h) { delete (HyperGraph*)h; }

// lambda-1, cut-net (over getPartitionsHypergraph(part,2,true)) and SOED, as printed by
// grading/grade.cpp:300-308.  Returns 0 on success, -1 on a thrown error.
int ref_metrics(void* h, const uint32_t* part, double* km1, double* cutnet, double* soed) {
  try {
    HyperGraph* hg = (HyperGraph*)h;
    std::vector<uint32_t> p(part, part + hg->nodes());
    HyperGraph q = hg->getPartitionsHypergraph(p, 2, true);
    *km1 = q.connectivity();
    *cutnet = q.cutnet();
    *soed = hg->soedFromPart(p);
    return 0;
  } catch (...) {
    return -1;
  }
}

// The -s and -p outputs (sources/runconfig.cpp:219-247): getPartitionsHypergraph(part,2,true)
// saved by savehMETIS / saveSNN / saveAXH as the path's extension says (which = 0), or the
// partition file by savePartitioning (which = 1).  Returns 0 on success, -1 on a thrown error.
int ref_save(void* h, const uint32_t* part, const char* path, int which) {
  try {
    HyperGraph* hg = (HyperGraph*)h;
    std::vector<uint32_t> p(part, part + hg->nodes());
    std::string s(path);
    if (which == 1) {
      HyperGraph::savePartitioning(s, p);
      return 0;
    }
    HyperGraph q = hg->getPartitionsHypergraph(p, 2, true);
    const std::string ext = s.size() > 4 ? s.substr(s.size() - 4) : "";
    if (ext == ".hgr") q.savehMETIS(s);
    else if (ext == ".snn") q.saveSNN(s);
    else if (ext == ".axh") q.saveAXH(s);
    else return -1;
    return 0;
  } catch (const std::exception& x) {
    fprintf(stderr, "ref_save: %s\n", x.what());
    return -1;
  }
}

// Constraints::checkPartitionValidity (sources/constr.cpp:90-147): 1 valid, 0 invalid,
// -1 when it throws (ids not sequential).
int ref_validity(void* h, const uint32_t* part, uint32_t omega, uint32_t delta, uint32_t max_parts) {
  try {
    HyperGraph* hg = (HyperGraph*)h;
    constraints::ConstraintsConfig c{"t", omega, delta, max_parts};
    constraints::Constraints con(c);
    std::vector<uint32_t> p(part, part + hg->nodes());
    return con.checkPartitionValidity(*hg, p, false) ? 1 : 0;
  } catch (...) {
    return -1;
  }
}
}
