// coarsen.hip - candidate proposal, deterministic pairing, orphan pairing, relabel.
//
// Candidates (reference candidatesProposal / candidates_kernel, kernels/coarsening.cu:10-213)
// are scored on the fly from the hypergraph instead of from materialised neighbour sets:
// each wave owns one node u and accumulates, in an LDS open-addressing table keyed by the
// neighbour v, score(u,v) = sum over shared hedges of floor(W_e/|e|) and the number of
// hedges inbound to both.  The pin stream is flattened across hedges (wave_iter.hpp) so
// small hedges do not idle the wave.  Nodes whose pin-visit count exceeds the LDS table get
// a workgroup and a global-memory table (batched to bound scratch).  Filtering (size cap,
// exact inbound-union cap), the symmetric noise and the top-`cnc` selection by
// (score desc, id desc) are exactly the reference's.
//
// Pairing (groupNodes, sources/coarsening.cu:70-210) replaces the reference's racing
// cooperative tree walk with an exact maximum-weight matching of each round's pointer
// forest, computed bottom-up by "last child continues" atomics (integer, order
// independent) and decided top-down along holder chains; see the oracle for the rule.
#include "prim.hpp"
#include "stages.hpp"
#include "wave_iter.hpp"

namespace hgp {

namespace {

constexpr uint32_t CAND_CAP = 1024;    // LDS slots per wave
constexpr uint32_t CAND_LOG = 10;
constexpr uint32_t CAND_LIGHT = 512;   // pin-visit bound served from LDS
constexpr uint32_t CAND_SLOTS = CAND_CAP / 64;
constexpr uint64_t HEAVY_BATCH_SLOTS = 64ull << 20;

__global__ void k_visit_bound(const uint64_t* __restrict__ hoff, const uint32_t* __restrict__ tch,
                              const uint64_t* __restrict__ toff, uint32_t n, uint64_t* __restrict__ bound,
                              uint32_t* __restrict__ light_flag) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  uint64_t b = 0;
  for (uint64_t j = toff[v]; j < toff[v + 1]; ++j) {
    uint32_t e = tch[j];
    b += hoff[e + 1] - hoff[e] - 1;
  }
  bound[v] = b;
  light_flag[v] = b <= CAND_LIGHT ? 1u : 0u;
}

__global__ void k_split_lists(const uint32_t* __restrict__ light_flag, const uint64_t* __restrict__ lpos, uint32_t n,
                              uint32_t* __restrict__ light, uint32_t* __restrict__ heavy) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  if (light_flag[v]) light[lpos[v]] = v;
  else heavy[v - lpos[v]] = v;
}

// candidate key = (final score << 32) | v, or 0 when v is filtered out
__device__ __forceinline__ uint64_t cand_key(uint32_t u, uint32_t v, uint32_t score, uint32_t inter, uint32_t size_u,
                                             uint32_t inb_u, const uint32_t* __restrict__ sizes,
                                             const uint32_t* __restrict__ inb, uint32_t omega, uint32_t delta) {
  if ((uint32_t)(size_u + sizes[v]) > omega) return 0;
  if ((uint32_t)((inb[v] - inter) + inb_u) > delta) return 0;
  const uint32_t s = score + noise64(u, v);
  if (s == 0) return 0;
  return ((uint64_t)s << 32) | v;
}

__device__ __forceinline__ void table_insert(uint32_t* keys, uint32_t* score, uint32_t* inter, uint32_t mask,
                                             uint32_t v, uint32_t contrib, bool shared_in) {
  uint32_t h = slot_hash(v) & mask;
  for (;;) {
    uint32_t old = atomicCAS(&keys[h], NONE, v);
    if (old == NONE || old == v) {
      atomicAdd(&score[h], contrib);
      if (shared_in) atomicAdd(&inter[h], 1u);
      return;
    }
    h = (h + 1) & mask;
  }
}

__global__ void __launch_bounds__(256) k_cand_light(const uint32_t* __restrict__ hedges, const uint64_t* __restrict__ hoff,
                                                    const uint32_t* __restrict__ srcc, const uint32_t* __restrict__ tch,
                                                    const uint64_t* __restrict__ toff, const uint32_t* __restrict__ inb,
                                                    const float* __restrict__ w, const uint32_t* __restrict__ sizes,
                                                    const uint32_t* __restrict__ list, uint32_t nlist, uint32_t cnc,
                                                    uint32_t omega, uint32_t delta, uint32_t* __restrict__ pairs,
                                                    uint32_t* __restrict__ scores) {
  __shared__ uint32_t s_keys[4][CAND_CAP];
  __shared__ uint32_t s_score[4][CAND_CAP];
  __shared__ uint32_t s_inter[4][CAND_CAP];
  const uint32_t wv = threadIdx.x / 64, l = lane_id();
  const uint32_t idx = blockIdx.x * 4 + wv;
  if (idx >= nlist) return;
  const uint32_t u = list[idx];
  uint32_t* keys = s_keys[wv];
  uint32_t* score = s_score[wv];
  uint32_t* inter = s_inter[wv];
  for (uint32_t s = l; s < CAND_CAP; s += 64) {
    keys[s] = NONE;
    score[s] = 0;
    inter[s] = 0;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const uint64_t t0 = toff[u], t1 = toff[u + 1];
  const uint32_t inb_u = inb[u];
  wave_visit_pins(tch, t0, t1, hoff, hedges, [&](const PinVisit& pv) {
    if (!pv.act || pv.pin == u) return;
    const uint32_t contrib = fixed_weight(w[pv.e]) / pv.len;
    const bool shared_in = (pv.j - t0) < inb_u && pv.pos >= srcc[pv.e];
    table_insert(keys, score, inter, CAND_CAP - 1, pv.pin, contrib, shared_in);
  });
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const uint32_t size_u = sizes[u];
  uint64_t cand[CAND_SLOTS];
#pragma unroll
  for (uint32_t s = 0; s < CAND_SLOTS; ++s) {
    const uint32_t slot = l + 64 * s;
    const uint32_t v = keys[slot];
    cand[s] = v == NONE ? 0ull : cand_key(u, v, score[slot], inter[slot], size_u, inb_u, sizes, inb, omega, delta);
  }
  for (uint32_t i = 0; i < cnc; ++i) {
    uint64_t m = 0;
#pragma unroll
    for (uint32_t s = 0; s < CAND_SLOTS; ++s) m = cand[s] > m ? cand[s] : m;
    const uint64_t best = wave_max(m);
#pragma unroll
    for (uint32_t s = 0; s < CAND_SLOTS; ++s)
      if (best != 0 && cand[s] == best) cand[s] = 0;
    if (l == 0) {
      pairs[(uint64_t)u * cnc + i] = best ? (uint32_t)best : NONE;
      scores[(uint64_t)u * cnc + i] = (uint32_t)(best >> 32);
    }
  }
}

// one workgroup per heavy node, table in global memory
__global__ void __launch_bounds__(256) k_cand_heavy(const uint32_t* __restrict__ hedges, const uint64_t* __restrict__ hoff,
                                                    const uint32_t* __restrict__ srcc, const uint32_t* __restrict__ tch,
                                                    const uint64_t* __restrict__ toff, const uint32_t* __restrict__ inb,
                                                    const float* __restrict__ w, const uint32_t* __restrict__ sizes,
                                                    const uint32_t* __restrict__ list, const uint64_t* __restrict__ tab_off,
                                                    uint32_t* __restrict__ gkeys, uint32_t* __restrict__ gscore,
                                                    uint32_t* __restrict__ ginter, uint32_t cnc, uint32_t omega,
                                                    uint32_t delta, uint32_t* __restrict__ pairs,
                                                    uint32_t* __restrict__ scores) {
  __shared__ uint64_t s_red[4];
  __shared__ uint64_t s_best;
  const uint32_t u = list[blockIdx.x];
  const uint64_t base = tab_off[blockIdx.x];
  const uint32_t cap = (uint32_t)(tab_off[blockIdx.x + 1] - base);
  uint32_t* keys = gkeys + base;
  uint32_t* score = gscore + base;
  uint32_t* inter = ginter + base;
  for (uint32_t s = threadIdx.x; s < cap; s += blockDim.x) {
    keys[s] = NONE;
    score[s] = 0;
    inter[s] = 0;
  }
  __threadfence_block();
  __syncthreads();
  const uint64_t t0 = toff[u], t1 = toff[u + 1];
  const uint32_t inb_u = inb[u];
  const uint32_t wv = threadIdx.x / 64;
  const uint64_t cnt = t1 - t0;
  const uint64_t a = t0 + cnt * wv / 4, b = t0 + cnt * (wv + 1) / 4;
  wave_visit_pins(tch, a, b, hoff, hedges, [&](const PinVisit& pv) {
    if (!pv.act || pv.pin == u) return;
    const uint32_t contrib = fixed_weight(w[pv.e]) / pv.len;
    const bool shared_in = (pv.j - t0) < inb_u && pv.pos >= srcc[pv.e];
    table_insert(keys, score, inter, cap - 1, pv.pin, contrib, shared_in);
  });
  __threadfence();
  __syncthreads();
  const uint32_t size_u = sizes[u];
  uint64_t prev = ~0ull;  // keys strictly below the previous winner remain eligible
  for (uint32_t i = 0; i < cnc; ++i) {
    uint64_t m = 0;
    for (uint32_t s = threadIdx.x; s < cap; s += blockDim.x) {
      const uint32_t v = keys[s];
      if (v == NONE) continue;
      const uint64_t k = cand_key(u, v, score[s], inter[s], size_u, inb_u, sizes, inb, omega, delta);
      if (k < prev && k > m) m = k;
    }
    m = wave_max(m);
    if (lane_id() == 0) s_red[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint64_t b2 = s_red[0];
      for (int q = 1; q < 4; ++q) b2 = s_red[q] > b2 ? s_red[q] : b2;
      s_best = b2;
      pairs[(uint64_t)u * cnc + i] = b2 ? (uint32_t)b2 : NONE;
      scores[(uint64_t)u * cnc + i] = (uint32_t)(b2 >> 32);
    }
    __syncthreads();
    prev = s_best;
    if (prev == 0) prev = 0;  // nothing left: later slots stay empty (k < 0 is never true)
  }
}

__global__ void k_heavy_caps(const uint32_t* __restrict__ heavy, const uint64_t* __restrict__ bound, uint32_t nh,
                             uint64_t* __restrict__ cap) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nh) return;
  uint64_t b = 2 * bound[heavy[i]] + 64;
  uint64_t c = 1;
  while (c < b) c <<= 1;
  cap[i] = c;
}

// ---------------------------------------------------------------------------------------
// pairing
__global__ void k_pair_init(const uint32_t* __restrict__ pairs, const uint32_t* __restrict__ scores, uint32_t cnc,
                            uint32_t round, uint32_t n, const uint32_t* __restrict__ raw, uint32_t* __restrict__ parent,
                            uint32_t* __restrict__ wgt, uint32_t* __restrict__ nchild, uint32_t* __restrict__ pending,
                            uint64_t* __restrict__ S, uint64_t* __restrict__ best, uint32_t* __restrict__ processed) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const bool locked = raw[v] != NONE;
  uint32_t par = NONE, wt = 0;
  if (!locked) {
    const uint32_t t = pairs[(uint64_t)v * cnc + round];
    if (t != NONE && raw[t] == NONE) {
      const uint32_t tt = pairs[(uint64_t)t * cnc + round];
      if (!(tt == v && v < t)) {  // a 2-cycle keeps the lower id as root
        par = t;
        wt = scores[(uint64_t)v * cnc + round];
      }
    }
  }
  parent[v] = par;
  wgt[v] = wt;
  nchild[v] = 0;
  pending[v] = 0;
  S[v] = 0;
  best[v] = 0;
  processed[v] = locked ? 1u : 0u;
}

__global__ void k_pair_count(const uint32_t* __restrict__ parent, uint32_t n, uint32_t* __restrict__ nchild,
                             uint32_t* __restrict__ pending) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint32_t p = parent[v];
  if (p != NONE) {
    atomicAdd(&nchild[p], 1u);
    atomicAdd(&pending[p], 1u);
  }
}

// Finish `cur` and keep climbing while this thread is the last child to arrive.
__device__ void sweep_up(uint32_t cur, const uint32_t* __restrict__ parent, const uint32_t* __restrict__ wgt,
                         uint32_t* pending, unsigned long long* S, unsigned long long* best, uint32_t* processed) {
  for (;;) {
    const uint64_t Sv = __hip_atomic_load(&S[cur], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t bv = __hip_atomic_load(&best[cur], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t bg = bv >> 32;
    const uint64_t wout = Sv + bg;
    __hip_atomic_store(&processed[cur], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t p = parent[cur];
    if (p == NONE) return;
    const int64_t gain = (int64_t)wgt[cur] - (int64_t)bg;
    const uint64_t key = ((uint64_t)(gain > 0 ? gain : 0) << 32) | cur;
    atomicAdd(&S[p], (unsigned long long)wout);
    atomicMax(&best[p], (unsigned long long)key);
    __threadfence();
    const uint32_t old = atomicSub(&pending[p], 1u);
    if (old != 1u) return;
    __threadfence();
    cur = p;
  }
}

__global__ void k_pair_sweep(const uint32_t* __restrict__ raw, const uint32_t* __restrict__ parent,
                             const uint32_t* __restrict__ wgt, const uint32_t* __restrict__ nchild, uint32_t n,
                             uint32_t* pending, unsigned long long* S, unsigned long long* best, uint32_t* processed) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  if (raw[v] != NONE || nchild[v] != 0) return;
  sweep_up(v, parent, wgt, pending, S, best, processed);
}

// nodes left unprocessed sit on cycles of the pointer graph: flag each cycle's minimum
__global__ void k_pair_cycles(const uint32_t* __restrict__ raw, const uint32_t* __restrict__ parent,
                              const uint32_t* __restrict__ processed, uint32_t n, uint32_t* __restrict__ cycmin) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  uint32_t flag = 0;
  if (raw[v] == NONE && !processed[v] && parent[v] != NONE) {
    uint32_t m = v, x = parent[v];
    for (uint32_t steps = 0; x != v && x != NONE && steps < n; ++steps) {
      m = x < m ? x : m;
      x = parent[x];
    }
    flag = (m == v) ? 1u : 0u;
  }
  cycmin[v] = flag;
}

__global__ void k_pair_break(const uint32_t* __restrict__ cycmin, uint32_t* parent, const uint32_t* __restrict__ wgt,
                             uint32_t n, uint32_t* pending, unsigned long long* S, unsigned long long* best,
                             uint32_t* processed) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n || !cycmin[v]) return;
  const uint32_t p = parent[v];
  parent[v] = NONE;
  __threadfence();
  const uint32_t old = atomicSub(&pending[p], 1u);
  if (old == 1u) {
    __threadfence();
    sweep_up(p, parent, wgt, pending, S, best, processed);
  }
}

__device__ __forceinline__ uint32_t holder_of(const unsigned long long* best, uint32_t v) {
  const uint64_t b = best[v];
  return (b >> 32) > 0 ? (uint32_t)b : NONE;
}

__global__ void k_pair_decide(const uint32_t* __restrict__ parent, const unsigned long long* __restrict__ best,
                              const uint32_t* __restrict__ locked0, uint32_t n, uint32_t* __restrict__ raw) {
  uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= n || locked0[y]) return;
  const uint32_t p = parent[y];
  if (p != NONE && holder_of(best, p) == y) return;  // y hangs below its parent's chain
  uint32_t cur = y;
  while (cur != NONE) {
    const uint32_t h = holder_of(best, cur);
    if (h == NONE) break;
    raw[cur] = h;
    raw[h] = h;
    cur = holder_of(best, h);
  }
}

__global__ void k_locked_snapshot(const uint32_t* __restrict__ raw, uint32_t n, uint32_t* __restrict__ locked0) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) locked0[v] = raw[v] != NONE;
}

// orphans
__global__ void k_orphan_flags(const uint32_t* __restrict__ pairs, uint32_t cnc, uint32_t n, uint32_t* __restrict__ f) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) f[v] = pairs[(uint64_t)v * cnc] == NONE;
}
__global__ void k_orphan_keys(const uint32_t* __restrict__ f, const uint64_t* __restrict__ pos,
                              const uint32_t* __restrict__ sizes, uint32_t n, uint64_t* __restrict__ keys) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n && f[v]) keys[pos[v]] = ((uint64_t)sizes[v] << 32) | v;
}
__global__ void k_orphan_pair(const uint64_t* __restrict__ keys, uint64_t m, const uint32_t* __restrict__ sizes,
                              const uint32_t* __restrict__ inb, uint32_t omega, uint32_t delta, uint32_t* __restrict__ raw) {
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m / 2) return;
  const uint32_t L = (uint32_t)keys[k], R = (uint32_t)keys[m - 1 - k];
  if ((uint32_t)(sizes[L] + sizes[R]) > omega) return;
  if ((uint32_t)(inb[L] + inb[R]) > delta) return;
  const uint32_t g = L < R ? L : R;
  raw[L] = g;
  raw[R] = g;
}

// relabel
__global__ void k_self_if_none(uint32_t* __restrict__ raw, uint32_t n) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n && raw[v] == NONE) raw[v] = v;
}
__global__ void k_mark(const uint32_t* __restrict__ raw, uint32_t n, uint32_t* __restrict__ flag) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) flag[raw[v]] = 1u;
}
__global__ void k_apply_ids(const uint32_t* __restrict__ raw, const uint64_t* __restrict__ nid,
                            const uint32_t* __restrict__ sizes, uint32_t n, uint32_t* __restrict__ groups,
                            uint32_t* __restrict__ gsizes) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint32_t g = (uint32_t)nid[raw[v]];
  groups[v] = g;
  atomicAdd(&gsizes[g], sizes[v]);
}

}  // namespace

void candidates(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const uint32_t* srcc, const uint32_t* tch,
                const uint64_t* toff, const uint32_t* inb, const float* w, const uint32_t* sizes, uint32_t n, uint32_t e,
                uint32_t cnc, uint32_t omega, uint32_t delta, uint32_t* pairs, uint32_t* scores) {
  (void)e;
  HGP_REQUIRE(cnc >= 1 && cnc <= 16, HGP_E_INVALID, "candidates: candidates_count must be in 1..16");
  if (n == 0) return;
  DBuf<uint64_t> bound(c, n);
  DBuf<uint32_t> lflag(c, n);
  HGP_LAUNCH(k_visit_bound, ceil_div(n, 256), 256, 0, c, hoff, tch, toff, n, bound.get(), lflag.get());
  DBuf<uint64_t> lpos(c, (uint64_t)n + 1);
  exclusive_scan(c, lflag.get(), lpos.get(), n);
  const uint32_t nl = (uint32_t)read_scalar(c, lpos.get() + n);
  const uint32_t nh = n - nl;
  DBuf<uint32_t> light(c, nl), heavy(c, nh);
  HGP_LAUNCH(k_split_lists, ceil_div(n, 256), 256, 0, c, lflag.get(), lpos.get(), n, light.get(), heavy.get());
  HGP_LAUNCH(k_cand_light, ceil_div(nl, 4), 256, 0, c, hedges, hoff, srcc, tch, toff, inb, w, sizes, light.get(), nl, cnc,
             omega, delta, pairs, scores);
  if (nh == 0) return;
  DBuf<uint64_t> caps(c, (uint64_t)nh + 1);
  HGP_LAUNCH(k_heavy_caps, ceil_div(nh, 256), 256, 0, c, heavy.get(), bound.get(), nh, caps.get());
  std::vector<uint64_t> hcap(nh);
  copy_d2h(c, hcap.data(), caps.get(), nh * sizeof(uint64_t));
  // batch heavy nodes so that their tables fit the scratch budget
  uint32_t b0 = 0;
  while (b0 < nh) {
    uint64_t tot = 0;
    uint32_t b1 = b0;
    while (b1 < nh && (b1 == b0 || tot + hcap[b1] <= HEAVY_BATCH_SLOTS)) tot += hcap[b1++];
    std::vector<uint64_t> off(b1 - b0 + 1, 0);
    for (uint32_t i = b0; i < b1; ++i) off[i - b0 + 1] = off[i - b0] + hcap[i];
    DBuf<uint64_t> doff(c, off.size());
    copy_h2d(c, doff.get(), off.data(), off.size() * sizeof(uint64_t));
    DBuf<uint32_t> gk(c, tot), gs(c, tot), gi(c, tot);
    HGP_LAUNCH(k_cand_heavy, b1 - b0, 256, 0, c, hedges, hoff, srcc, tch, toff, inb, w, sizes, heavy.get() + b0,
               doff.get(), gk.get(), gs.get(), gi.get(), cnc, omega, delta, pairs, scores);
    sync(c);  // host vector `off` must outlive the copy
    b0 = b1;
  }
}

void group_nodes(Ctx& c, const uint32_t* inb, const uint32_t* pairs, const uint32_t* scores, const uint32_t* sizes,
                 uint32_t n, uint32_t cnc, uint32_t omega, uint32_t delta, uint32_t& new_n, uint32_t*& groups_out,
                 uint32_t*& gsizes_out, uint32_t** ungroups_out, uint64_t** uoff_out, uint32_t* raw_out) {
  const uint32_t g = ceil_div(n, 256);
  DBuf<uint32_t> raw(c, n), parent(c, n), wgt(c, n), nchild(c, n), pending(c, n), processed(c, n), cycmin(c, n),
      locked0(c, n);
  DBuf<unsigned long long> S(c, n), best(c, n);
  fill_u32(c, raw.get(), NONE, n);
  for (uint32_t i = 0; i < cnc; ++i) {
    HGP_LAUNCH(k_pair_init, g, 256, 0, c, pairs, scores, cnc, i, n, raw.get(), parent.get(), wgt.get(), nchild.get(),
               pending.get(), (uint64_t*)S.get(), (uint64_t*)best.get(), processed.get());
    HGP_LAUNCH(k_pair_count, g, 256, 0, c, parent.get(), n, nchild.get(), pending.get());
    HGP_LAUNCH(k_pair_sweep, g, 256, 0, c, raw.get(), parent.get(), wgt.get(), nchild.get(), n, pending.get(), S.get(),
               best.get(), processed.get());
    HGP_LAUNCH(k_pair_cycles, g, 256, 0, c, raw.get(), parent.get(), processed.get(), n, cycmin.get());
    HGP_LAUNCH(k_pair_break, g, 256, 0, c, cycmin.get(), parent.get(), wgt.get(), n, pending.get(), S.get(), best.get(),
               processed.get());
    HGP_LAUNCH(k_locked_snapshot, g, 256, 0, c, raw.get(), n, locked0.get());
    HGP_LAUNCH(k_pair_decide, g, 256, 0, c, parent.get(), best.get(), locked0.get(), n, raw.get());
  }
  // orphans
  {
    DBuf<uint32_t> f(c, n);
    DBuf<uint64_t> pos(c, (uint64_t)n + 1);
    HGP_LAUNCH(k_orphan_flags, g, 256, 0, c, pairs, cnc, n, f.get());
    exclusive_scan(c, f.get(), pos.get(), n);
    const uint64_t m = read_scalar(c, pos.get() + n);
    if (m >= 2) {
      DBuf<uint64_t> keys(c, m);
      HGP_LAUNCH(k_orphan_keys, g, 256, 0, c, f.get(), pos.get(), sizes, n, keys.get());
      radix_sort(c, keys.get(), nullptr, m, 32 + std::max(1, bit_width(omega)));
      HGP_LAUNCH(k_orphan_pair, ceil_div(m / 2, 256), 256, 0, c, keys.get(), m, sizes, inb, omega, delta, raw.get());
    }
  }
  HGP_LAUNCH(k_self_if_none, g, 256, 0, c, raw.get(), n);
  if (raw_out) copy_d2h(c, raw_out, raw.get(), (uint64_t)n * sizeof(uint32_t));
  // relabel zero-based in ascending raw-id order
  DBuf<uint32_t> flag(c, n);
  memset_async(c, flag.get(), 0, (uint64_t)n * sizeof(uint32_t));
  HGP_LAUNCH(k_mark, g, 256, 0, c, raw.get(), n, flag.get());
  DBuf<uint64_t> nid(c, (uint64_t)n + 1);
  exclusive_scan(c, flag.get(), nid.get(), n);
  new_n = (uint32_t)read_scalar(c, nid.get() + n);
  DBuf<uint32_t> groups(c, n), gsizes(c, std::max<uint32_t>(new_n, 1));
  memset_async(c, gsizes.get(), 0, (uint64_t)new_n * sizeof(uint32_t));
  HGP_LAUNCH(k_apply_ids, g, 256, 0, c, raw.get(), nid.get(), sizes, n, groups.get(), gsizes.get());
  if (ungroups_out) {
    DBuf<uint32_t> keys(c, n), vals(c, n);
    copy_d2d(c, keys.get(), groups.get(), (uint64_t)n * sizeof(uint32_t));
    iota_u32(c, vals.get(), n);
    radix_sort(c, keys.get(), vals.get(), n, std::max(1, bit_width(new_n)));
    DBuf<uint64_t> uo(c, (uint64_t)new_n + 1);
    segment_offsets(c, keys.get(), 0, n, new_n, uo.get());
    *ungroups_out = vals.release();
    if (uoff_out) *uoff_out = uo.release();
  }
  groups_out = groups.release();
  gsizes_out = gsizes.release();
}

}  // namespace hgp
