// initpart.hip - a device initial partitioner for k-way (opt-in alternative to the host
// recursive bisection).
//
// Several starts, each a best-fit-decreasing packing of the coarsest level's nodes into k
// bins of capacity Omega (a seeded shuffle, then node sizes descending; ties in slack go to
// the lower bin), improved on the device by Jacobi label-propagation rounds until a round
// applies nothing.  The start with the lowest lambda-1 (then the lowest start index) is
// kept, so the result is a pure function of the input.  Modelled on the reference's legacy
// initial partitioner: host BFD packing (sources/init_part.cu:37-180) and Jacobi rounds of
// fm_refinement_gains_kernel + jacobi_apply_moves_kernel (:380-474); moves that lose are
// not applied here, and ties are broken deterministically instead of by coin flips.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <memory>
#include <numeric>
#include <vector>

#include "host/ip_core.hpp"
#include "prim.hpp"
#include "stages.hpp"

namespace hgp {

namespace {

// Best-fit decreasing in two levels: every node goes to the fullest bin it fits in below the
// average load min(Omega, ceil(W / k)) (W = total size), else to the fullest bin it fits in
// below Omega, else to the lightest bin (ties: lower bin).  Packing against the average load
// spreads the nodes over all k bins; packing against Omega alone fills bins to capacity and,
// with a few percent of slack, leaves the last ones empty.
void bfd_pack(const std::vector<uint32_t>& sizes, uint32_t k, uint32_t omega, uint64_t seed, std::vector<uint32_t>& part,
              std::vector<uint32_t>& loads) {
  const uint32_t n = (uint32_t)sizes.size();
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
  for (uint32_t i = n; i > 1; --i) {  // Fisher-Yates with a splitmix64 stream
    x = hgp_ip::splitmix64(x);
    std::swap(order[i - 1], order[x % i]);
  }
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return sizes[a] > sizes[b]; });
  part.assign(n, 0);
  loads.assign(k, 0);
  uint64_t total = 0;
  for (uint32_t s : sizes) total += s;
  const uint64_t avg = std::min<uint64_t>(omega, (total + k - 1) / k);
  for (uint32_t v : order) {
    const uint32_t s = sizes[v];
    uint32_t best = k;
    for (uint32_t b = 0; b < k; ++b)  // fullest bin below the average load
      if ((uint64_t)loads[b] + s <= avg && (best == k || loads[b] > loads[best])) best = b;
    if (best == k)
      for (uint32_t b = 0; b < k; ++b)  // fullest bin below Omega
        if ((uint64_t)loads[b] + s <= omega && (best == k || loads[b] > loads[best])) best = b;
    if (best == k) {  // lightest bin
      best = 0;
      for (uint32_t b = 1; b < k; ++b)
        if (loads[b] < loads[best]) best = b;
    }
    part[v] = best;
    loads[best] += s;
  }
}

// Multistart packing + Jacobi rounds on one level; the start with the lowest lambda-1 wins.
void multistart_lp(Ctx& c, const RefineLevel& L, uint32_t omega, uint32_t starts, uint32_t rounds, DBuf<uint32_t>& bp,
                   DBuf<uint32_t>& bps) {
  const uint32_t n = L.n, k = std::max<uint32_t>(L.k, 1);
  std::vector<uint32_t> sizes(n);
  if (n) copy_d2h(c, sizes.data(), L.sizes, (uint64_t)n * 4);
  bp = DBuf<uint32_t>(c, std::max<uint32_t>(n, 1));
  bps = DBuf<uint32_t>(c, k);
  DBuf<uint32_t> part(c, std::max<uint32_t>(n, 1)), ps(c, k);
  int64_t best_km1 = INT64_MAX;
  std::vector<uint32_t> hp, hl;
  for (uint32_t st = 0; st < std::max<uint32_t>(starts, 1); ++st) {
    bfd_pack(sizes, k, omega, st, hp, hl);
    copy_h2d(c, part.get(), hp.data(), (uint64_t)n * 4);
    copy_h2d(c, ps.get(), hl.data(), (uint64_t)k * 4);
    for (uint32_t r = 0; r < rounds; ++r)
      if (jacobi_round(c, L, omega, r, part.get(), ps.get()) == 0) break;
    const int64_t km1 = connectivity_fixed(c, L.hedges, L.hoff, L.w, part.get(), L.e);
    if (km1 < best_km1) {
      best_km1 = km1;
      copy_d2d(c, bp.get(), part.get(), (uint64_t)n * 4);
      copy_d2d(c, bps.get(), ps.get(), (uint64_t)k * 4);
    }
  }
}

}  // namespace

// The level's hedges are merged first (identical pin sets summed, single pins dropped: the
// cut is the same), then the level is coarsened further with the pipeline's own stages
// (candidates, pairing, contraction, merging again) down to about opt.nodes_per_part nodes
// per part; the multistart packing and label propagation run there, and the result is
// projected back level by level with the pipeline's k-way refinement.  The given level
// itself is left to the caller's refinement.
void initial_partition_device(Ctx& c, const RefineLevel& L0, uint32_t omega, const DeviceIpOptions& opt,
                              uint32_t prune, uint32_t*& part_out, uint32_t*& psizes_out) {
  struct Mini {
    RefineLevel L;
    DBuf<uint32_t> hedges, srcc, tch, inb, sizes, groups;
    DBuf<uint64_t> hoff, toff;
    DBuf<float> w;
    uint64_t P = 0;
  };
  using clk = std::chrono::high_resolution_clock;
  auto ms_since = [](clk::time_point t) { return std::chrono::duration<double, std::milli>(clk::now() - t).count(); };
  const auto t0 = clk::now();
  const uint32_t k = std::max<uint32_t>(L0.k, 1);
  const uint32_t target = std::max<uint32_t>(std::max<uint32_t>(opt.nodes_per_part, 1) * k, 256);
  const uint32_t cnc = 16, delta = 0x7FFFFFFFu;
  // a merged level of n nodes from (hedges, hoff, w) with e hedges; takes ownership of sizes
  auto make_level = [&](const uint32_t* hedges, const uint64_t* hoff, const float* w, uint32_t n, uint32_t e,
                        uint32_t prune_level, DBuf<uint32_t>&& sizes) {
    auto M = std::make_unique<Mini>();
    uint32_t me = 0, *mh;
    uint64_t* mo;
    float* mw;
    merge_level_device(c, hedges, hoff, w, n, e, prune_level, me, mh, mo, mw);
    M->hedges.ctx = &c, M->hedges.ptr = mh, M->hedges.n = 1;
    M->hoff.ctx = &c, M->hoff.ptr = mo, M->hoff.n = (uint64_t)me + 1;
    M->w.ctx = &c, M->w.ptr = mw, M->w.n = std::max<uint32_t>(me, 1);
    M->P = read_scalar(c, mo + me);
    M->srcc = DBuf<uint32_t>(c, std::max<uint32_t>(me, 1));
    memset_async(c, M->srcc.get(), 0, (uint64_t)me * 4);
    uint64_t ts = 0;
    uint32_t *tch, *inb;
    uint64_t* toff;
    build_touching(c, mh, mo, M->srcc.get(), n, me, ts, tch, toff, inb);
    M->tch.ctx = &c, M->tch.ptr = tch, M->tch.n = ts;
    M->toff.ctx = &c, M->toff.ptr = toff, M->toff.n = (uint64_t)n + 1;
    M->inb.ctx = &c, M->inb.ptr = inb, M->inb.n = n;
    M->sizes = std::move(sizes);
    M->L = RefineLevel{n, me, k, mh, mo, M->srcc.get(), tch, toff, inb, mw, M->sizes.get()};
    return M;
  };
  std::vector<std::unique_ptr<Mini>> lv;
  {
    DBuf<uint32_t> s0(c, std::max<uint32_t>(L0.n, 1));
    copy_d2d(c, s0.get(), L0.sizes, (uint64_t)L0.n * 4);
    lv.push_back(make_level(L0.hedges, L0.hoff, L0.w, L0.n, L0.e, prune, std::move(s0)));
  }
  if (c.verbose >= 2)
    fprintf(stderr, "  device initial partition: merged level %u nodes, %u hedges, %llu pins (%.1f ms)\n", L0.n,
            lv[0]->L.e, (unsigned long long)lv[0]->P, ms_since(t0));
  while (lv.back()->L.n > target && lv.size() < 32) {
    Mini& M = *lv.back();
    const RefineLevel& L = M.L;
    DBuf<uint32_t> pairs(c, (uint64_t)L.n * cnc), scores(c, (uint64_t)L.n * cnc);
    candidates(c, L.hedges, L.hoff, L.srcc, L.tch, L.toff, L.inb, L.w, L.sizes, L.n, L.e, cnc, omega, delta,
               pairs.get(), scores.get());
    uint32_t nn = 0;
    uint32_t *gp, *gs;
    group_nodes(c, L.inb, pairs.get(), scores.get(), L.sizes, L.n, cnc, omega, delta, nn, gp, gs, nullptr, nullptr);
    DBuf<uint32_t> groups, gsizes;
    groups.ctx = &c, groups.ptr = gp, groups.n = L.n;
    gsizes.ctx = &c, gsizes.ptr = gs, gsizes.n = std::max<uint32_t>(nn, 1);
    if ((float)nn / (float)L.n > 0.95f) break;
    uint64_t cs = 0;
    uint32_t *ch, *cc;
    uint64_t* co;
    coarsen_hedges(c, L.hedges, L.hoff, L.srcc, gp, L.e, M.P, nn, cs, ch, co, cc);
    DBuf<uint32_t> dch, dcc;
    DBuf<uint64_t> dco;
    dch.ctx = &c, dch.ptr = ch, dch.n = std::max<uint64_t>(cs, 1);
    dco.ctx = &c, dco.ptr = co, dco.n = (uint64_t)L.e + 1;
    dcc.ctx = &c, dcc.ptr = cc, dcc.n = std::max<uint32_t>(L.e, 1);
    auto N = make_level(ch, co, L.w, nn, L.e, 0u, std::move(gsizes));
    M.groups = std::move(groups);
    lv.push_back(std::move(N));
    if (c.verbose >= 2)
      fprintf(stderr, "  device initial partition: coarsened to %u nodes, %u hedges, %llu pins (%.1f ms)\n", nn,
              lv.back()->L.e, (unsigned long long)lv.back()->P, ms_since(t0));
  }
  DBuf<uint32_t> part, ps;
  multistart_lp(c, lv.back()->L, omega, opt.starts, opt.rounds, part, ps);
  if (c.verbose >= 2) {
    sync(c);
    fprintf(stderr, "  device initial partition: label propagation done (%.1f ms)\n", ms_since(t0));
  }
  for (size_t d = lv.size() - 1; d-- > 0;) {
    const RefineLevel& L = lv[d]->L;
    DBuf<uint32_t> fine(c, std::max<uint32_t>(L.n, 1));
    uncoarsen(c, lv[d]->groups.get(), part.get(), L.n, fine.get());
    part = std::move(fine);
    lv[d + 1].reset();
    refine(c, L, HGP_MODE_KWAY, omega, delta, opt.refine_repeats, false, part.get(), ps.get(), nullptr);
    if (c.verbose >= 2) {
      sync(c);
      fprintf(stderr, "  device initial partition: projected to %u nodes (%.1f ms)\n", L.n, ms_since(t0));
    }
  }
  sync(c);
  part_out = part.release();
  psizes_out = ps.release();
}

}  // namespace hgp
