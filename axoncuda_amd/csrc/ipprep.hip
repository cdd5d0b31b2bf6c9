// ipprep.hip - the initial partitioner's input built on the device.
//
// The coarsest level is handed to the initial partitioner the way the reference hands it to
// Mt-KaHyPar (sources/init_part.cu:733-953): hedges with at least two distinct pins, pins
// sorted, identical pin sets merged into the first of them with their 2^18 fixed-point
// weights summed, unique hedges in order of first appearance, node sizes as vertex weights.
// The host version of this step (hgp_ip::from_level) walks 6.7M pins on the c2 benchmark;
// here it is a handful of sorts and scans, and the host receives the merged hypergraph and
// its incidence lists ready to use.
//
// Before the merge, hedges are pruned as the reference prunes them (armonic_degree_score_kernel
// and prune_hedges_kernel, kernels/init_part.cu:221-281, launched at sources/init_part.cu:749-817):
// a hedge's score is its weight over the sum of its pins' inverse degrees, the expected number
// of kept hedges is f * N * floor(P / E) (f = 1 is the reference's target; 0 turns pruning
// off), and kept weights are divided by the keep probability.
// The reference draws with cuRAND Philox in float; this build uses exact integers, a
// splitmix64 draw per hedge and three IEEE double operations, so that the device and the
// oracle keep the same hedges with the same weights:
//   inv_e   = sum_{v in e} floor(2^24 / deg(v))                       (u64)
//   ratio_e = min(floor(W_e * 2^24 / inv_e), floor((2^64 - 1) / E))    (u64; 0 when inv_e = 0)
//   S = sum_e ratio_e,  T = f * N * floor(P / E)
//   p_e  = S == 0 ? 1 : min(1, (double)T * (double)ratio_e / (double)S)
//   keep = (double)(splitmix64(86 ^ e * 0x9E3779B97F4A7C15) >> 11) < p_e * 2^53
//   W'_e = keep ? min((uint64)((double)W_e / p_e), 2^40) : 0
#include <vector>

#include "host/ip_core.hpp"
#include "prim.hpp"
#include "stages.hpp"

namespace hgp {
namespace {

// (hedge << pb) | pin for every pin: one sort orders the pins inside every hedge
__global__ void k_ipp_keys(const uint32_t* __restrict__ hedges, const uint64_t* __restrict__ hoff, uint32_t e, int pb,
                           uint64_t* __restrict__ keys) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= e) return;
  for (uint64_t i = hoff[x]; i < hoff[x + 1]; ++i) keys[i] = ((uint64_t)x << pb) | hedges[i];
}

// keep[i] = the pin differs from its predecessor in the same hedge (drops repeated pins)
__global__ void k_ipp_keep(const uint64_t* __restrict__ keys, uint64_t p, uint32_t* __restrict__ keep) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p) return;
  keep[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// distinct pins of hedge x = kpos[hoff[x+1]] - kpos[hoff[x]]; compacted pin list
__global__ void k_ipp_compact(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ keep,
                              const uint64_t* __restrict__ kpos, uint64_t p, uint64_t pmask,
                              uint32_t* __restrict__ pins) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p || !keep[i]) return;
  pins[kpos[i]] = (uint32_t)(keys[i] & pmask);
}

// per hedge: distinct-pin range, hash of the sorted set, fixed-point weight; hedges of fewer
// than two distinct pins are flagged out
__global__ void k_ipp_hash(const uint64_t* __restrict__ hoff, const uint64_t* __restrict__ kpos,
                           const uint32_t* __restrict__ pins, const float* __restrict__ w,
                           const unsigned long long* __restrict__ pw, uint32_t e, uint64_t* __restrict__ hs,
                           uint64_t* __restrict__ start, uint32_t* __restrict__ len,
                           unsigned long long* __restrict__ wfx, uint32_t* __restrict__ live) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= e) return;
  const uint64_t s = kpos[hoff[x]], t = kpos[hoff[x + 1]];
  const uint32_t l = (uint32_t)(t - s);
  uint64_t h = 1469598103934665603ull;
  for (uint64_t i = s; i < t; ++i) {
    uint64_t z = (h ^ pins[i]) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    h = z ^ (z >> 31);
  }
  hs[x] = h;
  start[x] = s;
  len[x] = l;
  if (pw) {  // pruned: weight 0 marks a dropped hedge
    wfx[x] = pw[x];
    live[x] = l >= 2 && pw[x] != 0 ? 1u : 0u;
  } else {
    wfx[x] = fixed_weight(w[x]);
    live[x] = l >= 2 ? 1u : 0u;
  }
}

// deg[v] = number of pins equal to v (the touching degree)
__global__ void k_ipp_deg(const uint32_t* __restrict__ hedges, uint64_t p, uint32_t* __restrict__ deg) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < p) atomicAdd(&deg[hedges[i]], 1u);
}

// per hedge: ratio_e (see the header); the workgroup's sum goes to one of 64 spread counters
__global__ __launch_bounds__(256) void k_ipp_ratio(const uint32_t* __restrict__ hedges,
                                                   const uint64_t* __restrict__ hoff, const float* __restrict__ w,
                                                   const uint32_t* __restrict__ deg, uint32_t e, uint64_t cap,
                                                   unsigned long long* __restrict__ ratio,
                                                   unsigned long long* __restrict__ sums) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t r = 0;
  if (x < e) {
    uint64_t inv = 0;
    for (uint64_t i = hoff[x]; i < hoff[x + 1]; ++i) inv += (uint64_t)(16777216u / deg[hedges[i]]);
    if (inv) {
      r = ((uint64_t)fixed_weight(w[x]) << 24) / inv;
      if (r > cap) r = cap;
    }
    ratio[x] = r;
  }
  __shared__ uint64_t part[4];
  const uint64_t ws = wave_sum(r);
  if (lane_id() == 0) part[threadIdx.x / 64] = ws;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&sums[blockIdx.x % 64], (unsigned long long)(part[0] + part[1] + part[2] + part[3]));
}

__device__ __forceinline__ uint64_t splitmix64_dev(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// keep draw and scaled weight per hedge (see the header)
__global__ void k_ipp_prune(const float* __restrict__ w, const unsigned long long* __restrict__ ratio, uint32_t e,
                            double t, double sum, unsigned long long* __restrict__ pw) {
#pragma clang fp contract(off)
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= e) return;
  const double p = sum == 0.0 ? 1.0 : fmin(1.0, t * (double)ratio[x] / sum);
  const uint64_t u = splitmix64_dev(86ull ^ ((uint64_t)x * 0x9E3779B97F4A7C15ull)) >> 11;
  uint64_t out = 0;
  if ((double)u < p * 9007199254740992.0) {
    const double sc = (double)fixed_weight(w[x]) / p;
    out = sc >= 1099511627776.0 ? 1099511627776ull : (uint64_t)sc;
  }
  pw[x] = out;
}

__global__ void k_ipp_live_keys(const uint32_t* __restrict__ live, const uint64_t* __restrict__ lpos,
                                const uint64_t* __restrict__ hs, uint32_t e, uint64_t* __restrict__ keys,
                                uint32_t* __restrict__ ids) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= e || !live[x]) return;
  keys[lpos[x]] = hs[x];
  ids[lpos[x]] = x;
}

__device__ __forceinline__ bool same_set(const uint32_t* __restrict__ pins, const uint64_t* __restrict__ start,
                                         const uint32_t* __restrict__ len, uint32_t a, uint32_t b) {
  if (len[a] != len[b]) return false;
  const uint32_t* pa = pins + start[a];
  const uint32_t* pb = pins + start[b];
  for (uint32_t i = 0; i < len[a]; ++i)
    if (pa[i] != pb[i]) return false;
  return true;
}

// Sorted by (hash, hedge id): the first hedge of an identical set is the first identical
// entry of its hash run.  Runs hold one set unless hashes collide; then every earlier
// entry of the run is compared.
__global__ void k_ipp_rep(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ ids, uint32_t m,
                          const uint32_t* __restrict__ pins, const uint64_t* __restrict__ start,
                          const uint32_t* __restrict__ len, uint32_t* __restrict__ rep) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint64_t h = keys[i];
  const uint32_t x = ids[i];
  uint32_t s = i;
  while (s > 0 && keys[s - 1] == h) --s;
  uint32_t r = x;
  if (s < i) {
    if (same_set(pins, start, len, ids[s], x)) {
      r = ids[s];
    } else {
      for (uint32_t j = s + 1; j < i; ++j)
        if (same_set(pins, start, len, ids[j], x)) {
          r = ids[j];
          break;
        }
    }
  }
  rep[x] = r;
}

__global__ void k_ipp_first(const uint32_t* __restrict__ rep, uint32_t e, uint32_t* __restrict__ first) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < e) first[x] = rep[x] == x ? 1u : 0u;
}

// output length and weight per unique hedge (u = its rank among the first hedges)
__global__ void k_ipp_out_meta(const uint32_t* __restrict__ rep, const uint64_t* __restrict__ uid, uint32_t e,
                               const uint32_t* __restrict__ len, const unsigned long long* __restrict__ wfx,
                               uint64_t* __restrict__ olen, unsigned long long* __restrict__ ow) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= e || rep[x] == NONE) return;
  const uint32_t r = rep[x];
  const uint64_t u = uid[r];
  if (r == x) olen[u] = len[x];
  atomicAdd(&ow[u], wfx[x]);
}

// pins of the unique hedges at their output offsets
__global__ void k_ipp_out_pins(const uint32_t* __restrict__ rep, const uint64_t* __restrict__ uid, uint32_t e,
                               const uint32_t* __restrict__ pins, const uint64_t* __restrict__ start,
                               const uint32_t* __restrict__ len, const uint64_t* __restrict__ ooff,
                               uint32_t* __restrict__ opins) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= e || rep[x] != x) return;
  const uint64_t u = uid[x];
  const uint64_t o = ooff[u];
  for (uint32_t i = 0; i < len[x]; ++i) opins[o + i] = pins[start[x] + i];
}

// (pin, hedge) incidence keys of the merged hypergraph
__global__ void k_ipp_inc_keys(const uint64_t* __restrict__ ooff, const uint32_t* __restrict__ opins, uint32_t u,
                               int hb, uint64_t* __restrict__ ikeys) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= u) return;
  for (uint64_t i = ooff[x]; i < ooff[x + 1]; ++i) ikeys[i] = ((uint64_t)opins[i] << hb) | x;
}

__global__ void k_ipp_inc(const uint64_t* __restrict__ ikeys, uint64_t p, uint64_t hmask, uint32_t* __restrict__ inc) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < p) inc[i] = (uint32_t)(ikeys[i] & hmask);
}

}  // namespace

namespace {

__global__ void k_ipp_float_w(const unsigned long long* __restrict__ ow, uint32_t m, float* __restrict__ wf) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) wf[i] = (float)((double)ow[i] / 262144.0);
}

// Steps 1-3: sorted distinct pins, representatives, unique hedges in order of first
// appearance with their pins (opins, ooff) and summed fixed-point weights (ow); u = count.
struct Merged {
  DBuf<uint32_t> opins;
  DBuf<uint64_t> ooff;
  DBuf<unsigned long long> ow;
  uint32_t u = 0;
  uint64_t op = 0;
};

// Keep draws and scaled weights of the reference's hedge pruning (see the header): pw[e],
// 0 for a dropped hedge.
void prune_hedges(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, uint32_t n, uint32_t e,
                  uint64_t p, uint32_t mult, DBuf<unsigned long long>& pw) {
  DBuf<uint32_t> deg(c, n);
  memset_async(c, deg.get(), 0, (uint64_t)n * 4);
  HGP_LAUNCH(k_ipp_deg, ceil_div(p, 256), 256, 0, c, hedges, p, deg.get());
  DBuf<unsigned long long> ratio(c, e), sums(c, 64);
  memset_async(c, sums.get(), 0, 64 * 8);
  const uint64_t cap = ~0ull / e;
  HGP_LAUNCH(k_ipp_ratio, ceil_div(e, 256), 256, 0, c, hedges, hoff, w, deg.get(), e, cap, ratio.get(), sums.get());
  unsigned long long part[64];
  copy_d2h(c, part, sums.get(), sizeof part);
  uint64_t sum = 0;
  for (auto v : part) sum += v;
  const uint64_t t = (uint64_t)mult * n * (p / e);
  pw = DBuf<unsigned long long>(c, e);
  HGP_LAUNCH(k_ipp_prune, ceil_div(e, 256), 256, 0, c, w, ratio.get(), e, (double)t, (double)sum, pw.get());
}

void merge_hedges(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, uint32_t n, uint32_t e,
                  uint32_t prune, Merged& out) {
  out = Merged();
  if (e == 0 || n == 0) return;
  const uint64_t p = read_scalar(c, hoff + e);
  DBuf<unsigned long long> pw;
  if (prune) prune_hedges(c, hedges, hoff, w, n, e, p, prune, pw);
  const int pb = std::max(1, bit_width(n - 1));
  const int eb = std::max(1, bit_width(e - 1));
  HGP_REQUIRE(pb + eb <= 64, HGP_E_CAPACITY, "hedge merge: hedge and node ids exceed 64 key bits");
  const uint32_t ge = ceil_div(e, 256);
  // 1. sorted distinct pins per hedge
  DBuf<uint64_t> keys(c, std::max<uint64_t>(p, 1));
  HGP_LAUNCH(k_ipp_keys, ge, 256, 0, c, hedges, hoff, e, pb, keys.get());
  radix_sort(c, keys.get(), nullptr, p, pb + eb);
  DBuf<uint32_t> keep(c, std::max<uint64_t>(p, 1));
  DBuf<uint64_t> kpos(c, p + 1);
  HGP_LAUNCH(k_ipp_keep, ceil_div(p, 256), 256, 0, c, keys.get(), p, keep.get());
  exclusive_scan(c, keep.get(), kpos.get(), p);
  DBuf<uint32_t> pins(c, std::max<uint64_t>(p, 1));
  HGP_LAUNCH(k_ipp_compact, ceil_div(p, 256), 256, 0, c, keys.get(), keep.get(), kpos.get(), p,
             ((uint64_t)1 << pb) - 1, pins.get());
  keep.reset();
  keys.reset();
  // 2. per hedge: hash, range, weight; hedges of >= 2 pins sorted by (hash, id)
  DBuf<uint64_t> hs(c, e), start(c, e);
  DBuf<uint32_t> len(c, e), live(c, e);
  DBuf<unsigned long long> wfx(c, e);
  HGP_LAUNCH(k_ipp_hash, ge, 256, 0, c, hoff, kpos.get(), pins.get(), w, pw.get(), e, hs.get(), start.get(), len.get(),
             wfx.get(), live.get());
  pw.reset();
  DBuf<uint64_t> lpos(c, (uint64_t)e + 1);
  exclusive_scan(c, live.get(), lpos.get(), e);
  const uint32_t m = (uint32_t)read_scalar(c, lpos.get() + e);
  if (m == 0) return;
  DBuf<uint64_t> hk(c, m);
  DBuf<uint32_t> ids(c, m);
  HGP_LAUNCH(k_ipp_live_keys, ge, 256, 0, c, live.get(), lpos.get(), hs.get(), e, hk.get(), ids.get());
  radix_sort(c, hk.get(), ids.get(), m, 64);
  // 3. representative of every hedge; the first hedges in id order are the output
  DBuf<uint32_t> rep(c, e), first(c, e);
  fill_u32(c, rep.get(), NONE, e);
  HGP_LAUNCH(k_ipp_rep, ceil_div(m, 256), 256, 0, c, hk.get(), ids.get(), m, pins.get(), start.get(), len.get(),
             rep.get());
  HGP_LAUNCH(k_ipp_first, ge, 256, 0, c, rep.get(), e, first.get());
  DBuf<uint64_t> uid(c, (uint64_t)e + 1);
  exclusive_scan(c, first.get(), uid.get(), e);
  const uint32_t u = (uint32_t)read_scalar(c, uid.get() + e);
  DBuf<uint64_t> olen(c, u);
  out.ooff = DBuf<uint64_t>(c, (uint64_t)u + 1);
  out.ow = DBuf<unsigned long long>(c, u);
  memset_async(c, out.ow.get(), 0, (uint64_t)u * 8);
  HGP_LAUNCH(k_ipp_out_meta, ge, 256, 0, c, rep.get(), uid.get(), e, len.get(), wfx.get(), olen.get(), out.ow.get());
  exclusive_scan(c, olen.get(), out.ooff.get(), u);
  out.op = read_scalar(c, out.ooff.get() + u);
  out.u = u;
  out.opins = DBuf<uint32_t>(c, out.op);
  HGP_LAUNCH(k_ipp_out_pins, ge, 256, 0, c, rep.get(), uid.get(), e, pins.get(), start.get(), len.get(),
             out.ooff.get(), out.opins.get());
}

}  // namespace

void merge_level_device(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, uint32_t n, uint32_t e,
                        uint32_t prune, uint32_t& e_out, uint32_t*& hedges_out, uint64_t*& hoff_out, float*& w_out) {
  Merged M;
  merge_hedges(c, hedges, hoff, w, n, e, prune, M);
  e_out = M.u;
  if (M.u == 0) {
    M.ooff = DBuf<uint64_t>(c, 1);
    memset_async(c, M.ooff.get(), 0, 8);
  }
  DBuf<float> wf(c, std::max<uint32_t>(M.u, 1));
  if (M.u) HGP_LAUNCH(k_ipp_float_w, ceil_div(M.u, 256), 256, 0, c, M.ow.get(), M.u, wf.get());
  if (!M.opins.get()) M.opins = DBuf<uint32_t>(c, 1);
  hedges_out = M.opins.release();
  hoff_out = M.ooff.release();
  w_out = wf.release();
}

void ip_input_device(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, const uint32_t* sizes,
                     uint32_t n, uint32_t e, uint32_t prune, hgp_ip::Hypergraph& h) {
  h = hgp_ip::Hypergraph();
  h.n = n;
  h.vw.resize(n);
  if (n) {
    std::vector<uint32_t> s(n);
    copy_d2h(c, s.data(), sizes, (uint64_t)n * 4);
    for (uint32_t v = 0; v < n; ++v) h.vw[v] = s[v];
  }
  h.off.assign(1, 0);
  h.ioff.assign((uint64_t)n + 1, 0);
  Merged M;
  merge_hedges(c, hedges, hoff, w, n, e, prune, M);
  const uint32_t u = M.u;
  if (u == 0) return;
  const uint64_t op = M.op;
  const int pb = std::max(1, bit_width(n - 1));
  const int hb = std::max(1, bit_width(u - 1));
  HGP_REQUIRE(pb + hb <= 64, HGP_E_CAPACITY, "initial partition input: hedge and node ids exceed 64 key bits");
  // 4. incidence: hedges of every node in ascending order
  DBuf<uint64_t> ikeys(c, op);
  HGP_LAUNCH(k_ipp_inc_keys, ceil_div(u, 256), 256, 0, c, M.ooff.get(), M.opins.get(), u, hb, ikeys.get());
  radix_sort(c, ikeys.get(), nullptr, op, pb + hb);
  DBuf<uint32_t> inc(c, op);
  DBuf<uint64_t> ioff(c, (uint64_t)n + 1);
  HGP_LAUNCH(k_ipp_inc, ceil_div(op, 256), 256, 0, c, ikeys.get(), op, ((uint64_t)1 << hb) - 1, inc.get());
  segment_offsets(c, ikeys.get(), hb, op, n, ioff.get());
  h.off.resize((uint64_t)u + 1);
  h.w.resize(u);
  h.pins.resize(op);
  h.inc.resize(op);
  copy_d2h(c, h.off.data(), M.ooff.get(), ((uint64_t)u + 1) * 8);
  copy_d2h(c, h.w.data(), M.ow.get(), (uint64_t)u * 8);
  copy_d2h(c, h.pins.data(), M.opins.get(), op * 4);
  copy_d2h(c, h.inc.data(), inc.get(), op * 4);
  copy_d2h(c, h.ioff.data(), ioff.get(), ((uint64_t)n + 1) * 8);
}

}  // namespace hgp
