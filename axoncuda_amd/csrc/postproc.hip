// postproc.hip - greedy merges and relabels (reference sources/postprocess.cu, sources/main.cu:817-830).
//
// Both greedyMergeGroups (-ipm, :16-245) and mergeSmallPartitions (:247-420) pack items
// sorted by (size, inbound, id) into consecutive segments whose summed size and summed
// inbound count stay within the caps.  The segment starting at position i ends where either
// prefix sum would first exceed its cap (an upper_bound on the prefix arrays); the packs are
// then numbered by binary lifting over that monotone "next start" relation, which keeps the
// whole packing parallel (log2(m) gather rounds) - the same scheme as the reference.
#include <vector>

#include "prim.hpp"
#include "stages.hpp"

namespace hgp {

namespace {

__device__ __forceinline__ uint64_t upper_bound_u64(const uint64_t* a, uint64_t n, uint64_t x) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t m = (lo + hi) / 2;
    if (a[m] <= x) lo = m + 1;
    else hi = m;
  }
  return lo;
}

__global__ void k_next_start(const uint64_t* __restrict__ ps, const uint64_t* __restrict__ pi, uint32_t m,
                             uint64_t omega, uint64_t delta, uint32_t* __restrict__ ns) {
  uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos > m) return;
  if (pos == m) {
    ns[m] = m;
    return;
  }
  const uint64_t us = upper_bound_u64(ps, (uint64_t)m + 1, ps[pos] + omega);
  const uint64_t ui = upper_bound_u64(pi, (uint64_t)m + 1, pi[pos] + delta);
  const uint64_t t = (us < ui ? us : ui) - 1;
  ns[pos] = (uint32_t)(t > pos + 1 ? t : pos + 1);
}

__global__ void k_jump(const uint32_t* __restrict__ j0, uint32_t m, uint32_t* __restrict__ j1) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= m) j1[i] = j0[j0[i]];
}

__global__ void k_pack_ids(const uint32_t* __restrict__ jumps, uint32_t levels, uint32_t m,
                           uint32_t* __restrict__ pack) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  uint32_t cur = 0, id = 0;
  for (int k = (int)levels - 1; k >= 0; --k) {
    const uint32_t cand = jumps[(uint64_t)k * (m + 1) + cur];
    if (cand <= i) {
      cur = cand;
      id += 1u << k;
    }
  }
  pack[i] = id;
}

__global__ void k_group_inbound(const uint32_t* __restrict__ groups, const uint32_t* __restrict__ inb, uint32_t n,
                                uint32_t* gin) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) atomicAdd(&gin[groups[v]], inb[v]);
}

__global__ void k_sort_keys_si(const uint32_t* __restrict__ s, const uint32_t* __restrict__ in,
                               const uint32_t* __restrict__ ids, uint32_t m, uint64_t* __restrict__ keys,
                               uint32_t* __restrict__ vals) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t id = ids ? ids[i] : i;
  keys[i] = ((uint64_t)s[id] << 32) | in[id];
  vals[i] = id;
}

__global__ void k_sorted_values(const uint64_t* __restrict__ keys, uint32_t m, uint64_t* __restrict__ sz,
                                uint64_t* __restrict__ in) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  sz[i] = keys[i] >> 32;
  in[i] = keys[i] & 0xFFFFFFFFull;
}

__global__ void k_rep_min(const uint32_t* __restrict__ pack, const uint32_t* __restrict__ ids, uint32_t m,
                          uint32_t* rep) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) atomicMin(&rep[pack[i]], ids[i]);
}

__global__ void k_map_scatter(const uint32_t* __restrict__ pack, const uint32_t* __restrict__ ids,
                              const uint32_t* __restrict__ rep, uint32_t m, uint32_t* __restrict__ map) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) map[ids[i]] = rep[pack[i]];
}

__global__ void k_remap(uint32_t* __restrict__ x, const uint32_t* __restrict__ map, uint32_t n) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) x[v] = map[x[v]];
}

__global__ void k_mark_ids(const uint32_t* __restrict__ x, uint32_t n, uint32_t* __restrict__ flag) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) flag[x[v]] = 1u;
}

__global__ void k_renumber(uint32_t* __restrict__ x, const uint64_t* __restrict__ nid, uint32_t n) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) x[v] = (uint32_t)nid[x[v]];
}

__global__ void k_sizes_of(const uint32_t* __restrict__ x, const uint32_t* __restrict__ sizes, uint32_t n,
                           uint32_t* out) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) atomicAdd(&out[x[v]], sizes[v]);
}

__global__ void k_small_flags(const uint32_t* __restrict__ psizes, uint32_t k, uint32_t* __restrict__ f) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < k) f[p] = psizes[p] < 16u ? 1u : 0u;  // SMALL_PART_MERGE_SIZE_THRESHOLD (headers/defines.cuh:33)
}

__global__ void k_compact_ids(const uint32_t* __restrict__ f, const uint64_t* __restrict__ pos, uint32_t k,
                              uint32_t* __restrict__ out) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < k && f[p]) out[pos[p]] = p;
}

__global__ void k_max_id(const uint32_t* __restrict__ x, uint32_t n, uint32_t* out) {
  uint32_t m = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = x[i] > m ? x[i] : m;
  m = wave_max(m);
  if (lane_id() == 0) atomicMax(out, m);
}

// packs[m] for items already sorted; sizes/inbounds as u64 per sorted position
void greedy_pack(Ctx& c, const uint64_t* ssz, const uint64_t* sin, uint32_t m, uint64_t omega, uint64_t delta,
                 uint32_t* pack) {
  DBuf<uint64_t> ps(c, (uint64_t)m + 1), pi(c, (uint64_t)m + 1);
  exclusive_scan(c, ssz, ps.get(), m);
  exclusive_scan(c, sin, pi.get(), m);
  uint32_t levels = 1;
  while ((1ull << (levels - 1)) < m) levels++;
  DBuf<uint32_t> jumps(c, (uint64_t)levels * (m + 1));
  HGP_LAUNCH(k_next_start, ceil_div((uint64_t)m + 1, 256), 256, 0, c, ps.get(), pi.get(), m, omega, delta, jumps.get());
  for (uint32_t k = 1; k < levels; ++k)
    HGP_LAUNCH(k_jump, ceil_div((uint64_t)m + 1, 256), 256, 0, c, jumps.get() + (uint64_t)(k - 1) * (m + 1), m,
               jumps.get() + (uint64_t)k * (m + 1));
  HGP_LAUNCH(k_pack_ids, ceil_div(m, 256), 256, 0, c, jumps.get(), levels, m, pack);
}

uint32_t renumber(Ctx& c, uint32_t* x, uint32_t n, uint32_t idspace) {
  DBuf<uint32_t> flag(c, idspace);
  memset_async(c, flag.get(), 0, (uint64_t)idspace * 4);
  HGP_LAUNCH(k_mark_ids, ceil_div(n, 256), 256, 0, c, x, n, flag.get());
  DBuf<uint64_t> nid(c, (uint64_t)idspace + 1);
  exclusive_scan(c, flag.get(), nid.get(), idspace);
  HGP_LAUNCH(k_renumber, ceil_div(n, 256), 256, 0, c, x, nid.get(), n);
  return (uint32_t)read_scalar(c, nid.get() + idspace);
}

}  // namespace

void greedy_merge_groups(Ctx& c, const uint32_t* sizes, const uint32_t* inb, uint32_t n, uint32_t ng, uint32_t omega,
                         uint32_t delta, uint32_t* groups, uint32_t* gsizes, uint32_t& new_ng) {
  new_ng = ng;
  if (n == 0 || ng < 2) return;
  DBuf<uint32_t> gin(c, ng);
  memset_async(c, gin.get(), 0, (uint64_t)ng * 4);
  HGP_LAUNCH(k_group_inbound, ceil_div(n, 256), 256, 0, c, groups, inb, n, gin.get());
  DBuf<uint64_t> keys(c, ng);
  DBuf<uint32_t> ids(c, ng);
  HGP_LAUNCH(k_sort_keys_si, ceil_div(ng, 256), 256, 0, c, gsizes, gin.get(), nullptr, ng, keys.get(), ids.get());
  radix_sort(c, keys.get(), ids.get(), ng, 64);
  DBuf<uint64_t> ssz(c, ng), sin(c, ng);
  HGP_LAUNCH(k_sorted_values, ceil_div(ng, 256), 256, 0, c, keys.get(), ng, ssz.get(), sin.get());
  DBuf<uint32_t> pack(c, ng), rep(c, ng), gmap(c, ng);
  greedy_pack(c, ssz.get(), sin.get(), ng, omega, delta, pack.get());
  fill_u32(c, rep.get(), NONE, ng);
  HGP_LAUNCH(k_rep_min, ceil_div(ng, 256), 256, 0, c, pack.get(), ids.get(), ng, rep.get());
  HGP_LAUNCH(k_map_scatter, ceil_div(ng, 256), 256, 0, c, pack.get(), ids.get(), rep.get(), ng, gmap.get());
  HGP_LAUNCH(k_remap, ceil_div(n, 256), 256, 0, c, groups, gmap.get(), n);
  new_ng = renumber(c, groups, n, ng);
  memset_async(c, gsizes, 0, (uint64_t)new_ng * 4);
  HGP_LAUNCH(k_sizes_of, ceil_div(n, 256), 256, 0, c, groups, sizes, n, gsizes);
}

void merge_small_partitions(Ctx& c, const uint32_t* psizes, const uint32_t* pinbound, uint32_t n, uint32_t k,
                            uint32_t omega, uint32_t delta, uint32_t* part) {
  if (k == 0) return;
  DBuf<uint32_t> f(c, k);
  DBuf<uint64_t> pos(c, (uint64_t)k + 1);
  HGP_LAUNCH(k_small_flags, ceil_div(k, 256), 256, 0, c, psizes, k, f.get());
  exclusive_scan(c, f.get(), pos.get(), k);
  const uint32_t m = (uint32_t)read_scalar(c, pos.get() + k);
  if (m == 0) return;
  DBuf<uint32_t> small(c, m), ids(c, m);
  HGP_LAUNCH(k_compact_ids, ceil_div(k, 256), 256, 0, c, f.get(), pos.get(), k, small.get());
  DBuf<uint64_t> keys(c, m);
  HGP_LAUNCH(k_sort_keys_si, ceil_div(m, 256), 256, 0, c, psizes, pinbound, small.get(), m, keys.get(), ids.get());
  radix_sort(c, keys.get(), ids.get(), m, 64);
  DBuf<uint64_t> ssz(c, m), sin(c, m);
  HGP_LAUNCH(k_sorted_values, ceil_div(m, 256), 256, 0, c, keys.get(), m, ssz.get(), sin.get());
  DBuf<uint32_t> pack(c, m), rep(c, m), pmap(c, k);
  greedy_pack(c, ssz.get(), sin.get(), m, omega, delta, pack.get());
  fill_u32(c, rep.get(), NONE, m);
  HGP_LAUNCH(k_rep_min, ceil_div(m, 256), 256, 0, c, pack.get(), ids.get(), m, rep.get());
  iota_u32(c, pmap.get(), k);
  HGP_LAUNCH(k_map_scatter, ceil_div(m, 256), 256, 0, c, pack.get(), ids.get(), rep.get(), m, pmap.get());
  HGP_LAUNCH(k_remap, ceil_div(n, 256), 256, 0, c, part, pmap.get(), n);
}

void relabel(Ctx& c, uint32_t n, uint32_t* part, uint32_t& new_k) {
  if (n == 0) {
    new_k = 0;
    return;
  }
  DBuf<uint32_t> mx(c, 1);
  memset_async(c, mx.get(), 0, 4);
  HGP_LAUNCH(k_max_id, std::min<uint32_t>(ceil_div(n, 256), 2048), 256, 0, c, part, n, mx.get());
  const uint32_t space = read_scalar(c, mx.get()) + 1;
  new_k = renumber(c, part, n, space);
}

}  // namespace hgp
