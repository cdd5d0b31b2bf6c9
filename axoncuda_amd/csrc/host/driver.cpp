// driver.cpp - whole-pipeline orchestration (reference sources/main.cu:38-940).
//
// The recursion coarsen_refine_uncoarsen (:282-784) runs as an explicit stack of device
// levels: candidates -> pairing -> base cases (KWAY :358-391 / :448-485, INCC :495-564) ->
// contraction (coarse hedges, coarse touching) -> optional host spill (-smh, :631-706) ->
// next level; then uncoarsen + refine back up, INCC small-partition merge, final relabel.
// The KWAY base case calls the initial partitioner (Mt-KaHyPar when installed, else the
// built-in recursive bisection of ip_core.hpp).
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "../prim.hpp"
#include "../stages.hpp"
#include "driver.hpp"
#include "ip_core.hpp"

namespace hgp {

namespace {

std::string find_mtkahypar() {
  if (const char* env = std::getenv("HGP_MTKAHYPAR")) return env;  // explicit binary, "" disables
  const char* path = std::getenv("PATH");
  if (!path) return "";
  std::stringstream ss(path);
  std::string dir;
  while (std::getline(ss, dir, ':')) {
    std::string cand = dir + "/mtkahypar";
    if (::access(cand.c_str(), X_OK) == 0) return cand;
  }
  return "";
}

std::string float_shortest(float x) {
  char buf[64];
  for (int prec = 1; prec < 12; ++prec) {
    std::snprintf(buf, sizeof buf, "%.*g", prec, (double)x);
    if (std::strtof(buf, nullptr) == x) return buf;
  }
  std::snprintf(buf, sizeof buf, "%.9g", (double)x);
  return buf;
}

// Mt-KaHyPar as an external process, with the reference's file protocol and flags
// (sources/init_part.cu:821-938).  Returns false when the run or its output fails.
bool run_mtkahypar(const std::string& bin, const hgp_ip::Hypergraph& h, uint32_t k, float eps, uint32_t* part) {
  char tmpl[] = "/tmp/hgp_ip_XXXXXX";
  char* dir = ::mkdtemp(tmpl);
  if (!dir) return false;
  const std::string d(dir), file = d + "/coarse_tmp.hgr";
  {
    std::ofstream f(file);
    f << (h.off.size() - 1) << " " << h.n << " 11\n";
    for (size_t e = 0; e + 1 < h.off.size(); ++e) {
      f << (uint32_t)h.w[e];
      for (uint64_t i = h.off[e]; i < h.off[e + 1]; ++i) f << " " << (h.pins[i] + 1);
      f << "\n";
    }
    for (uint32_t v = 0; v < h.n; ++v) f << h.vw[v] << "\n";
  }
  const std::string es = float_shortest(eps);
  std::ostringstream cmd;
  cmd << bin << " -h " << file << " -k " << k << " -e " << es << " -t 16 -o km1 -v 0 -m direct"
      << " --write-partition-file 1 --partition-output-folder " << d << " --preset-type default --seed 86";
  const int rc = std::system(cmd.str().c_str());
  std::ostringstream out;
  out << file << ".part" << k << ".epsilon" << es << ".seed86.KaHyPar";
  bool ok = rc == 0;
  if (ok) {
    std::ifstream f(out.str());
    for (uint32_t v = 0; v < h.n && ok; ++v) ok = (bool)(f >> part[v]) && part[v] < k;
  }
  std::string rm = "rm -rf '" + d + "'";
  (void)std::system(rm.c_str());
  return ok;
}

}  // namespace

void initial_partition(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, uint32_t n, uint32_t e,
                       const uint32_t* sizes, uint32_t k, float eps, uint32_t omega, int32_t prune_arg,
                       uint32_t*& part_out, uint32_t*& psizes_out, double* ms, const RefineLevel* device_level,
                       const DeviceIpOptions& dip) {
  auto t0 = std::chrono::high_resolution_clock::now();
  static const std::string mtk = find_mtkahypar();
  const uint32_t prune = prune_arg >= 0 ? (uint32_t)prune_arg : (mtk.empty() ? 0u : 1u);
  if (device_level && mtk.empty()) {  // device multistart partitioner
    RefineLevel RL = *device_level;
    RL.k = k;
    initial_partition_device(c, RL, omega, dip, prune, part_out, psizes_out);
    if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
    return;
  }
  const auto t1 = t0;
  hgp_ip::Hypergraph h;
  ip_input_device(c, hedges, hoff, w, sizes, n, e, prune, h);
  const auto t2 = std::chrono::high_resolution_clock::now();
  std::vector<uint32_t> part(n, 0), psizes(k, 0);
  bool done = false;
  if (!mtk.empty()) {
    done = run_mtkahypar(mtk, h, k, eps, part.data());
    if (done) {
      for (uint32_t v = 0; v < n; ++v) psizes[part[v]] += (uint32_t)h.vw[v];
      for (uint32_t p = 0; p < k; ++p)
        if (psizes[p] > omega) done = false;  // the reference rejects oversized parts (init_part.cu:930-936)
    }
    if (!done) {
      std::fill(psizes.begin(), psizes.end(), 0u);
      if (c.verbose) fprintf(stderr, "WARNING: mtkahypar failed, using the built-in initial partitioner\n");
    }
  }
  if (!done) {
    try {
      hgp_ip::partition(h, k, eps, omega, part.data(), psizes.data());
    } catch (const hgp_ip::Infeasible& x) {
      throw Error(HGP_E_INFEASIBLE, x.what());
    }
  }
  const auto t3 = std::chrono::high_resolution_clock::now();
  if (c.verbose >= 2) {
    using msd = std::chrono::duration<double, std::milli>;
    fprintf(stderr, "  initial partition: input on the device %.1f ms (%zu hedges, %zu pins), partition %.1f ms\n",
            msd(t2 - t1).count(), h.off.size() - 1, h.pins.size(), msd(t3 - t2).count());
  }
  DBuf<uint32_t> dp(c, n), ds(c, k);
  copy_h2d(c, dp.get(), part.data(), (uint64_t)n * 4);
  copy_h2d(c, ds.get(), psizes.data(), (uint64_t)k * 4);
  sync(c);
  part_out = dp.release();
  psizes_out = ds.release();
  if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
}

namespace {

struct LevelDev {
  uint32_t n = 0;
  uint64_t P = 0;
  DBuf<uint32_t> hedges, srcc, tch, inb, sizes, groups;
  DBuf<uint64_t> hoff, toff;
  std::vector<uint32_t> h_hedges, h_srcc, h_tch, h_inb;
  std::vector<uint64_t> h_hoff, h_toff;
  bool spilled = false;
};

template <typename T>
void spill(Ctx& c, DBuf<T>& d, std::vector<T>& h) {
  h.resize(d.n);
  copy_d2h(c, h.data(), d.get(), d.n * sizeof(T));
  d.reset();
}
template <typename T>
void unspill(Ctx& c, DBuf<T>& d, std::vector<T>& h) {
  d = DBuf<T>(c, h.size());
  copy_h2d(c, d.get(), h.data(), h.size() * sizeof(T));
  sync(c);
  std::vector<T>().swap(h);
}

}  // namespace

void partition(Ctx& c, const hgp_params& prm, const PipelineInput& in, uint32_t* part_out, hgp_run_stats* stats) {
  const bool kw = prm.mode == HGP_MODE_KWAY;
  const uint32_t n0 = in.n, E = in.e;
  const uint64_t P0 = in.p;
  const uint32_t omega = prm.max_nodes_per_part, delta = prm.max_inbound_per_part, max_parts = prm.max_parts;
  HGP_REQUIRE(omega > 0 && delta > 0 && max_parts > 0, HGP_E_INVALID, "constraints must be positive");
  HGP_REQUIRE(prm.candidates_count >= 1 && prm.candidates_count <= 16, HGP_E_INVALID, "-cnc must be in 1..16");
  HGP_REQUIRE(!kw || prm.kway >= 1, HGP_E_INVALID, "-k must be positive");
  // K-way runs size the partition tables by k; the reference sets max_parts = k in that mode
  // (sources/runconfig.cpp:190-192), so any other value is a caller error.
  HGP_REQUIRE(!kw || max_parts == prm.kway, HGP_E_INVALID, "k-way mode needs max_parts == kway");
  const uint32_t cnc = prm.candidates_count;
  struct BudgetScope {  // the budget holds for this call only
    Ctx& c;
    uint64_t saved;
    ~BudgetScope() { c.event_budget = saved; }
  } budget_scope{c, c.event_budget};
  c.event_budget = prm.refine_event_budget;
  DeviceIpOptions dip;
  if (prm.ip_starts) dip.starts = prm.ip_starts;
  if (prm.ip_rounds) dip.rounds = prm.ip_rounds;
  if (prm.ip_nodes_per_part) dip.nodes_per_part = prm.ip_nodes_per_part;
  if (prm.ip_refine_repeats) dip.refine_repeats = prm.ip_refine_repeats;
  const uint32_t target_parts = std::min(max_parts, (uint32_t)(((uint64_t)n0 + omega - 1) / omega));
  hgp_run_stats st{};
  st.total_pins = P0;
  const uint64_t held_before = c.in_use;
  c.peak = c.in_use;

  hipEvent_t ev_start, ev_core, ev_stop;
  HGP_HIP(hipEventCreate(&ev_start));
  HGP_HIP(hipEventCreate(&ev_core));
  HGP_HIP(hipEventCreate(&ev_stop));
  HGP_HIP(hipEventRecord(ev_start, c.stream));

  std::vector<LevelDev> lv(1);
  DBuf<float> w(c, E);
  {
    LevelDev& L = lv[0];
    L.n = n0;
    L.P = P0;
    L.hedges = DBuf<uint32_t>(c, P0);
    L.hoff = DBuf<uint64_t>(c, (uint64_t)E + 1);
    L.srcc = DBuf<uint32_t>(c, E);
    L.sizes = DBuf<uint32_t>(c, n0);
    if (in.host) {
      copy_h2d(c, L.hedges.get(), in.host->pins.data(), P0 * 4);
      copy_h2d(c, L.hoff.get(), in.host->off.data(), ((uint64_t)E + 1) * 8);
      copy_h2d(c, L.srcc.get(), in.host->srcc.data(), (uint64_t)E * 4);
      copy_h2d(c, w.get(), in.host->w.data(), (uint64_t)E * 4);
    } else {
      copy_d2d(c, L.hedges.get(), in.d_hedges, P0 * 4);
      copy_d2d(c, L.hoff.get(), in.d_hoff, ((uint64_t)E + 1) * 8);
      copy_d2d(c, L.srcc.get(), in.d_srcc, (uint64_t)E * 4);
      copy_d2d(c, w.get(), in.d_w, (uint64_t)E * 4);
    }
    fill_u32(c, L.sizes.get(), 1u, n0);  // every input node counts 1 (sources/main.cu:217-218)
    HGP_HIP(hipEventRecord(ev_core, c.stream));
    uint64_t ts;
    uint32_t *tch, *inb;
    uint64_t* toff;
    build_touching(c, L.hedges.get(), L.hoff.get(), L.srcc.get(), n0, E, ts, tch, toff, inb);
    L.tch.ptr = tch, L.tch.ctx = &c, L.tch.n = ts;
    L.toff.ptr = toff, L.toff.ctx = &c, L.toff.n = (uint64_t)n0 + 1;
    L.inb.ptr = inb, L.inb.ctx = &c, L.inb.n = n0;
  }
  const double bytes_touching = 8.0 * P0 + 12.0 * E + 12.0 * n0;
  double bytes = bytes_touching;

  DBuf<uint32_t> part, psizes, pinb;
  uint32_t K = 0;
  uint32_t depth = 0;
  uint32_t repeats_total = 0;
  for (;; ++depth) {
    HGP_REQUIRE(depth < 63, HGP_E_INTERNAL, "too many coarsening levels");
    LevelDev& L = lv[depth];
    const uint32_t n = L.n;
    st.level_nodes[depth] = n;
    st.level_pins[depth] = L.P;
    if (c.verbose) fprintf(stderr, "Coarsening level %u, remaining nodes=%u, pins=%llu\n", depth, n, (unsigned long long)L.P);
    const RefineLevel ipl{L.n, E, prm.kway, L.hedges.get(), L.hoff.get(), L.srcc.get(), L.tch.get(), L.toff.get(),
                          L.inb.get(), w.get(), L.sizes.get()};
    const bool dev_ip = prm.initial_partitioner == HGP_IP_DEVICE ||
                        (prm.initial_partitioner == HGP_IP_AUTO && L.P >= HGP_IP_AUTO_DEVICE_PINS);
    const RefineLevel* ip_dev = dev_ip ? &ipl : nullptr;
    if (kw && n < 8192) {  // KWAY_INIT_UPPER_THREASHOLD (headers/defines.cuh:35)
      uint32_t *pp, *ps;
      initial_partition(c, L.hedges.get(), L.hoff.get(), w.get(), n, E, L.sizes.get(), prm.kway, prm.epsilon, omega,
                        prm.initial_prune, pp, ps, &st.ip_ms, ip_dev, dip);
      part = DBuf<uint32_t>();
      part.ctx = &c, part.ptr = pp, part.n = n;
      psizes.ctx = &c, psizes.ptr = ps, psizes.n = prm.kway;
      K = prm.kway;
      break;
    }
    DBuf<uint32_t> pairs(c, (uint64_t)n * cnc), scores(c, (uint64_t)n * cnc);
    candidates(c, L.hedges.get(), L.hoff.get(), L.srcc.get(), L.tch.get(), L.toff.get(), L.inb.get(), w.get(),
               L.sizes.get(), n, E, cnc, omega, delta, pairs.get(), scores.get());
    uint32_t nn = 0;
    uint32_t *gp, *gs;
    group_nodes(c, L.inb.get(), pairs.get(), scores.get(), L.sizes.get(), n, cnc, omega, delta, nn, gp, gs, nullptr,
                nullptr);
    DBuf<uint32_t> groups, gsizes;
    groups.ctx = &c, groups.ptr = gp, groups.n = n;
    gsizes.ctx = &c, gsizes.ptr = gs, gsizes.n = std::max<uint32_t>(nn, 1);
    bytes += 4.0 * L.P + 4.0 * L.P + 16.0 * E + 16.0 * n + 8.0 * cnc * n;  // candidates
    bytes += 8.0 * cnc * n + 16.0 * n;                                      // pairing + relabel
    const float ratio = (float)nn / (float)n;
    if (c.verbose) fprintf(stderr, "  grouped %u -> %u nodes (ratio %.3f)\n", n, nn, ratio);
    if (kw && (ratio > 0.95f || nn < 1024)) {  // KWAY_INIT_SHRINK_RATIO_LIMIT / LOWER_THREASHOLD
      uint32_t *pp, *ps;
      initial_partition(c, L.hedges.get(), L.hoff.get(), w.get(), n, E, L.sizes.get(), prm.kway, prm.epsilon, omega,
                        prm.initial_prune, pp, ps, &st.ip_ms, ip_dev, dip);
      part.ctx = &c, part.ptr = pp, part.n = n;
      psizes.ctx = &c, psizes.ptr = ps, psizes.n = prm.kway;
      K = prm.kway;
      break;
    }
    if (!kw && (nn <= target_parts || nn == n || (depth >= 32 && ratio > 0.95f))) {
      if (prm.initial_partitions_merge)
        greedy_merge_groups(c, L.sizes.get(), L.inb.get(), n, nn, omega, delta, groups.get(), gsizes.get(), nn);
      if (c.verbose) fprintf(stderr, "Initial partitioning built at level %u: %u partitions\n", depth, nn);
      part = std::move(groups);
      psizes = std::move(gsizes);
      K = nn;
      break;
    }
    LevelDev next;
    next.n = nn;
    {
      uint64_t cs;
      uint32_t *ch, *csc;
      uint64_t* cho;
      coarsen_hedges(c, L.hedges.get(), L.hoff.get(), L.srcc.get(), groups.get(), E, L.P, nn, cs, ch, cho, csc);
      next.P = cs;
      next.hedges.ctx = &c, next.hedges.ptr = ch, next.hedges.n = cs;
      next.hoff.ctx = &c, next.hoff.ptr = cho, next.hoff.n = (uint64_t)E + 1;
      next.srcc.ctx = &c, next.srcc.ptr = csc, next.srcc.n = E;
      uint64_t ts;
      uint32_t *tch, *inb;
      uint64_t* toff;
      build_touching(c, ch, cho, csc, nn, E, ts, tch, toff, inb);
      next.tch.ctx = &c, next.tch.ptr = tch, next.tch.n = ts;
      next.toff.ctx = &c, next.toff.ptr = toff, next.toff.n = (uint64_t)nn + 1;
      next.inb.ctx = &c, next.inb.ptr = inb, next.inb.n = nn;
    }
    next.sizes = std::move(gsizes);
    bytes += 2.0 * (4.0 * L.P + 8.0 * E) + 2.0 * (4.0 * L.P + 12.0 * n) + 4.0 * n;  // contraction
    L.groups = std::move(groups);
    if (depth < prm.save_memory_up_to_level) {  // -smh: park the fine level on the host
      spill(c, L.hedges, L.h_hedges);
      spill(c, L.hoff, L.h_hoff);
      spill(c, L.srcc, L.h_srcc);
      spill(c, L.tch, L.h_tch);
      spill(c, L.toff, L.h_toff);
      spill(c, L.inb, L.h_inb);
      L.spilled = true;
    }
    lv.push_back(std::move(next));
  }
  st.levels = depth + 1;

  pinb = DBuf<uint32_t>(c, std::max<uint32_t>(K, 1));
  memset_async(c, pinb.get(), 0, (uint64_t)K * 4);
  bool refined_level0 = false;
  for (int d = (int)depth - 1; d >= 0; --d) {
    LevelDev& L = lv[d];
    lv.resize(d + 1);  // free the coarser level
    if (L.spilled) {
      unspill(c, L.hedges, L.h_hedges);
      unspill(c, L.hoff, L.h_hoff);
      unspill(c, L.srcc, L.h_srcc);
      unspill(c, L.tch, L.h_tch);
      unspill(c, L.toff, L.h_toff);
      unspill(c, L.inb, L.h_inb);
      L.spilled = false;
    }
    DBuf<uint32_t> fine(c, L.n);
    uncoarsen(c, L.groups.get(), part.get(), L.n, fine.get());
    part = std::move(fine);
    L.groups.reset();
    RefineLevel RL{L.n, E, K, L.hedges.get(), L.hoff.get(), L.srcc.get(), L.tch.get(), L.toff.get(), L.inb.get(),
                   w.get(), L.sizes.get()};
    if (c.verbose) fprintf(stderr, "Uncoarsening level %d, nodes=%u\n", d, L.n);
    if (kw && prm.lp_rounds) lp_refine(c, RL, omega, prm.lp_rounds, part.get(), psizes.get());
    refine(c, RL, prm.mode, omega, delta, prm.refine_repeats, d == 0, part.get(), psizes.get(), pinb.get());
    repeats_total += prm.refine_repeats;
    bytes += prm.refine_repeats * (4.0 * L.P + 12.0 * E + 4.0 * L.P + 8.0 * L.n + 4.0 * L.n + 16.0 * L.n);
    if (d == 0) refined_level0 = true;
  }
  if (!kw) {
    LevelDev& L = lv[0];
    if (!refined_level0) inbound_sizes(c, L.hedges.get(), L.hoff.get(), L.srcc.get(), part.get(), E, K, pinb.get());
    merge_small_partitions(c, psizes.get(), pinb.get(), n0, K, omega, delta, part.get());
  }
  uint32_t nk = 0;
  relabel(c, n0, part.get(), nk);
  st.num_partitions = nk;
  HGP_HIP(hipEventRecord(ev_stop, c.stream));
  copy_d2h(c, part_out, part.get(), (uint64_t)n0 * 4);
  HGP_HIP(hipEventSynchronize(ev_stop));
  float core_ms = 0, dev_ms = 0;
  HGP_HIP(hipEventElapsedTime(&core_ms, ev_core, ev_stop));
  HGP_HIP(hipEventElapsedTime(&dev_ms, ev_start, ev_stop));
  HGP_HIP(hipEventDestroy(ev_start));
  HGP_HIP(hipEventDestroy(ev_core));
  HGP_HIP(hipEventDestroy(ev_stop));
  st.core_ms = core_ms;
  st.device_ms = dev_ms;
  st.bytes_algorithmic = (uint64_t)bytes;
  st.peak_device_bytes = c.peak - held_before;
  (void)repeats_total;
  if (stats) *stats = st;
}

}  // namespace hgp
