// construct.hip - incidence ("touching"), neighbour sets, coarse hedges, uncoarsening.
//
// Touching (reference buildTouching, sources/construction.cu:64-179) is built as a
// stable transpose: every pin becomes the key (node*2 + is_source) with its hedge id as
// value; a stable radix sort then yields, per node, the inbound hedges ascending followed
// by the outbound hedges ascending - the canonical order, independent of any atomic
// arrival order, in two coalesced passes per digit instead of per-node hashing.
// Coarse touching (coarsenTouching) is the same transpose applied to the coarse hedges.
//
// Coarse hedges (coarsenHedges, :563-676) and neighbour sets (buildNeighbors :242-340,
// coarsenNeighbors :342-561) are sort-unique problems: the (segment, value) pairs are
// radix-sorted on exactly the bits they need and run heads are compacted with scans.
#include "prim.hpp"
#include "stages.hpp"
#include "wave_iter.hpp"

namespace hgp {

namespace {

__global__ void k_touch_keys(const uint32_t* __restrict__ hedges, const uint64_t* __restrict__ hoff,
                             const uint32_t* __restrict__ srcc, uint32_t E, uint32_t* __restrict__ keys,
                             uint32_t* __restrict__ vals) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  uint64_t s = hoff[e], t = hoff[e + 1];
  uint32_t sc = srcc[e];
  for (uint64_t i = s; i < t; ++i) {
    keys[i] = hedges[i] * 2u + ((i - s) < sc ? 1u : 0u);
    vals[i] = e;
  }
}

// inbound_count[v]: number of even keys (destination incidences) in v's segment
__global__ void k_inbound_counts(const uint32_t* __restrict__ keys, const uint64_t* __restrict__ toff, uint32_t n,
                                 uint32_t* __restrict__ inb) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  uint64_t lo = toff[v], hi = toff[v + 1];
  const uint32_t odd = 2u * v + 1u;
  uint64_t a = lo, b = hi;  // first index with key >= odd
  while (a < b) {
    uint64_t m = (a + b) / 2;
    if (keys[m] < odd) a = m + 1;
    else b = m;
  }
  inb[v] = (uint32_t)(a - lo);
}

// per-node count of (pin, node) incidences excluding the node itself
__global__ void k_nbr_bound(const uint64_t* __restrict__ hoff, const uint32_t* __restrict__ tch,
                            const uint64_t* __restrict__ toff, uint32_t n, uint64_t* __restrict__ bound) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  uint64_t b = 0;
  for (uint64_t j = toff[v]; j < toff[v + 1]; ++j) {
    uint32_t e = tch[j];
    b += hoff[e + 1] - hoff[e] - 1;
  }
  bound[v] = b;
}

__global__ void __launch_bounds__(256) k_nbr_gen(const uint32_t* __restrict__ hedges, const uint64_t* __restrict__ hoff,
                                                 const uint32_t* __restrict__ tch, const uint64_t* __restrict__ toff,
                                                 uint32_t n, const uint64_t* __restrict__ outoff,
                                                 uint64_t* __restrict__ keys, int pb) {
  const uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
  if (v >= n) return;
  uint64_t pos = outoff[v];
  wave_visit_pins(tch, toff[v], toff[v + 1], hoff, hedges, [&](const PinVisit& pv) {
    const bool keep = pv.act && pv.pin != v;
    const uint64_t m = __ballot(keep);
    if (keep) keys[pos + lanes_below(m)] = ((uint64_t)v << pb) | pv.pin;
    pos += __popcll(m);
  });
}

// coarse neighbour pairs: (groups[m], groups[x]) for x in N(m); same-group pairs get a
// sentinel segment id (all ones) that sorts after every real group
__global__ void k_cnbr_gen(const uint32_t* __restrict__ nbr, const uint64_t* __restrict__ noff,
                           const uint32_t* __restrict__ groups, uint32_t n, uint64_t* __restrict__ keys, int pb) {
  const uint32_t m = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
  if (m >= n) return;
  const uint32_t g = groups[m];
  const uint64_t sentinel = ((uint64_t)1 << (2 * pb)) - 1;
  for (uint64_t i = noff[m] + lane_id(); i < noff[m + 1]; i += 64) {
    uint32_t h = groups[nbr[i]];
    keys[i] = h == g ? sentinel : (((uint64_t)g << pb) | h);
  }
}

template <typename K>
__global__ void k_head_flags(const K* __restrict__ keys, uint64_t n, int shift, uint32_t* __restrict__ flags) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = (i == 0 || (keys[i] >> shift) != (keys[i - 1] >> shift)) ? 1u : 0u;
}

__global__ void k_compact_heads(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flags,
                                const uint64_t* __restrict__ pos, uint64_t n, uint64_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flags[i]) return;
  out[pos[i]] = keys[i];
}

__global__ void k_low_bits(const uint64_t* __restrict__ keys, uint64_t n, uint64_t mask, uint32_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint32_t)(keys[i] & mask);
}

// coarse hedge keys: (e << (gb+1)) | (group << 1) | is_source
__global__ void k_chedge_keys(const uint32_t* __restrict__ hedges, const uint64_t* __restrict__ hoff,
                              const uint32_t* __restrict__ srcc, const uint32_t* __restrict__ groups, uint32_t E,
                              int gb, uint64_t* __restrict__ keys) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  uint64_t s = hoff[e], t = hoff[e + 1];
  uint32_t sc = srcc[e];
  for (uint64_t i = s; i < t; ++i)
    keys[i] = ((uint64_t)e << (gb + 1)) | ((uint64_t)groups[hedges[i]] << 1) | ((i - s) < sc ? 1u : 0u);
}

// run heads of (e, group): a head whose first entry is a destination is a destination
// group; otherwise every entry of the run is a source and it is a source group
__global__ void k_chedge_flags(const uint64_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ srcf,
                               uint32_t* __restrict__ dstf) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool head = i == 0 || (keys[i] >> 1) != (keys[i - 1] >> 1);
  bool is_src = keys[i] & 1u;
  srcf[i] = head && is_src;
  dstf[i] = head && !is_src;
}

__global__ void k_chedge_counts(const uint64_t* __restrict__ seg, const uint64_t* __restrict__ sr,
                                const uint64_t* __restrict__ dr, uint32_t E, uint64_t* __restrict__ cnt,
                                uint32_t* __restrict__ csrcc) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  uint64_t a = seg[e], b = seg[e + 1];
  uint64_t ns = sr[b] - sr[a], nd = dr[b] - dr[a];
  cnt[e] = ns + nd;
  csrcc[e] = (uint32_t)ns;
}

__global__ void k_chedge_write(const uint64_t* __restrict__ keys, uint64_t n, int gb, const uint64_t* __restrict__ seg,
                               const uint32_t* __restrict__ srcf, const uint32_t* __restrict__ dstf,
                               const uint64_t* __restrict__ sr, const uint64_t* __restrict__ dr,
                               const uint64_t* __restrict__ choff, const uint32_t* __restrict__ csrcc,
                               uint32_t* __restrict__ chedges) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!srcf[i] && !dstf[i]) return;
  const uint64_t k = keys[i];
  const uint32_t e = (uint32_t)(k >> (gb + 1));
  const uint32_t g = (uint32_t)((k >> 1) & (((uint64_t)1 << gb) - 1));
  const uint64_t a = seg[e];
  const uint32_t ns = csrcc[e];
  if (srcf[i]) {
    chedges[choff[e] + (sr[i] - sr[a])] = g;
  } else {
    const uint64_t nd = choff[e + 1] - choff[e] - ns;
    chedges[choff[e] + ns + (nd - 1 - (dr[i] - dr[a]))] = g;
  }
}

__global__ void k_max_group_of_pins(const uint32_t* __restrict__ hedges, const uint32_t* __restrict__ groups, uint64_t p,
                                    uint32_t* out) {
  uint32_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t g = groups[hedges[i]];
    m = g > m ? g : m;
  }
  m = wave_max(m);
  if (lane_id() == 0) atomicMax(out, m);
}

__global__ void k_uncoarsen(const uint32_t* __restrict__ groups, const uint32_t* __restrict__ cpart, uint32_t n,
                            uint32_t* __restrict__ part) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) part[v] = cpart[groups[v]];
}

// Sorted (segment << pb | value) keys with duplicates -> CSR of unique values per segment.
void unique_csr(Ctx& c, DBuf<uint64_t>& keys, uint64_t m, int pb, uint32_t nseg, uint64_t& out_size,
                uint32_t*& out_vals, uint64_t*& out_off) {
  DBuf<uint32_t> flags(c, m);
  DBuf<uint64_t> pos(c, m + 1);
  HGP_LAUNCH((k_head_flags<uint64_t>), ceil_div(m, 256), 256, 0, c, keys.get(), m, 0, flags.get());
  exclusive_scan(c, flags.get(), pos.get(), m);
  uint64_t u = m ? read_scalar(c, pos.get() + m) : 0;
  DBuf<uint64_t> ukeys(c, u);
  HGP_LAUNCH(k_compact_heads, ceil_div(m, 256), 256, 0, c, keys.get(), flags.get(), pos.get(), m, ukeys.get());
  flags.reset();
  pos.reset();
  keys.reset();
  DBuf<uint64_t> off(c, (uint64_t)nseg + 1);
  segment_offsets(c, ukeys.get(), pb, u, nseg, off.get());
  out_size = read_scalar(c, off.get() + nseg);  // entries of real segments (sentinels excluded)
  DBuf<uint32_t> vals(c, out_size);
  HGP_LAUNCH(k_low_bits, ceil_div(out_size, 256), 256, 0, c, ukeys.get(), out_size, (((uint64_t)1 << pb) - 1),
             vals.get());
  out_vals = vals.release();
  out_off = off.release();
}

}  // namespace

void build_touching(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const uint32_t* srcc, uint32_t n, uint32_t e,
                    uint64_t& tsize, uint32_t*& tch, uint64_t*& toff, uint32_t*& inb) {
  HGP_REQUIRE(n < 0x7FFFFFFFu, HGP_E_INVALID, "build_touching: too many nodes for 31-bit ids");
  const uint64_t p = e ? read_scalar(c, hoff + e) : 0;
  DBuf<uint32_t> keys(c, p);
  DBuf<uint32_t> vals(c, p);
  HGP_LAUNCH(k_touch_keys, ceil_div(e, 256), 256, 0, c, hedges, hoff, srcc, e, keys.get(), vals.get());
  radix_sort(c, keys.get(), vals.get(), p, bit_width(2ull * n + 1));
  DBuf<uint64_t> off(c, (uint64_t)n + 1);
  segment_offsets(c, keys.get(), 1, p, n, off.get());
  DBuf<uint32_t> ib(c, n);
  HGP_LAUNCH(k_inbound_counts, ceil_div(n, 256), 256, 0, c, keys.get(), off.get(), n, ib.get());
  tsize = p;
  tch = vals.release();
  toff = off.release();
  inb = ib.release();
}

void build_neighbors(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const uint32_t* tch, const uint64_t* toff,
                     uint32_t n, uint64_t& nsize, uint32_t*& nbr, uint64_t*& noff) {
  const int pb = std::max(1, bit_width(n));
  HGP_REQUIRE(2 * pb <= 64, HGP_E_INVALID, "build_neighbors: node ids too wide");
  DBuf<uint64_t> bound(c, (uint64_t)n + 1);
  HGP_LAUNCH(k_nbr_bound, ceil_div(n, 256), 256, 0, c, hoff, tch, toff, n, bound.get());
  exclusive_scan(c, bound.get(), bound.get(), n);
  const uint64_t m = read_scalar(c, bound.get() + n);
  DBuf<uint64_t> keys(c, m);
  HGP_LAUNCH(k_nbr_gen, ceil_div((uint64_t)n * 64, 256), 256, 0, c, hedges, hoff, tch, toff, n, bound.get(),
             keys.get(), pb);
  bound.reset();
  radix_sort(c, keys.get(), nullptr, m, 2 * pb);
  unique_csr(c, keys, m, pb, n, nsize, nbr, noff);
}

void coarsen_neighbors(Ctx& c, const uint32_t* nbr, const uint64_t* noff, const uint32_t* groups, uint32_t n,
                       uint32_t new_n, uint64_t& nsize, uint32_t*& cnbr, uint64_t*& cnoff) {
  const int pb = std::max(1, bit_width(new_n));  // 2^pb - 1 > every group id: sentinel space
  const uint64_t m = n ? read_scalar(c, noff + n) : 0;
  DBuf<uint64_t> keys(c, m);
  HGP_LAUNCH(k_cnbr_gen, ceil_div((uint64_t)n * 64, 256), 256, 0, c, nbr, noff, groups, n, keys.get(), pb);
  radix_sort(c, keys.get(), nullptr, m, 2 * pb);
  unique_csr(c, keys, m, pb, new_n, nsize, cnbr, cnoff);
}

void coarsen_hedges(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const uint32_t* srcc, const uint32_t* groups,
                    uint32_t e, uint64_t hsize, uint32_t new_n, uint64_t& csize, uint32_t*& chedges, uint64_t*& choff,
                    uint32_t*& csrcc) {
  if (new_n == 0 && hsize > 0) {  // derive the coarse node count from the groups of the pins
    DBuf<uint32_t> mx(c, 1);
    memset_async(c, mx.get(), 0, 4);
    HGP_LAUNCH(k_max_group_of_pins, std::min<uint32_t>(ceil_div(hsize, 256), 2048), 256, 0, c, hedges, groups, hsize,
               mx.get());
    new_n = read_scalar(c, mx.get()) + 1;
  }
  const int gb = std::max(1, bit_width(new_n ? new_n - 1 : 0));
  const int eb = std::max(1, bit_width(e ? e - 1 : 0));
  HGP_REQUIRE(eb + gb + 1 <= 64, HGP_E_CAPACITY, "coarsen_hedges: hedge and group ids exceed 64 key bits");
  const uint64_t p = hsize;
  DBuf<uint64_t> keys(c, p);
  HGP_LAUNCH(k_chedge_keys, ceil_div(e, 256), 256, 0, c, hedges, hoff, srcc, groups, e, gb, keys.get());
  radix_sort(c, keys.get(), nullptr, p, eb + gb + 1);
  DBuf<uint32_t> srcf(c, p), dstf(c, p);
  HGP_LAUNCH(k_chedge_flags, ceil_div(p, 256), 256, 0, c, keys.get(), p, srcf.get(), dstf.get());
  DBuf<uint64_t> sr(c, p + 1), dr(c, p + 1);
  exclusive_scan(c, srcf.get(), sr.get(), p);
  exclusive_scan(c, dstf.get(), dr.get(), p);
  DBuf<uint64_t> seg(c, (uint64_t)e + 1);
  segment_offsets(c, keys.get(), gb + 1, p, e, seg.get());
  DBuf<uint64_t> off(c, (uint64_t)e + 1);
  DBuf<uint32_t> sc(c, e);
  HGP_LAUNCH(k_chedge_counts, ceil_div(e, 256), 256, 0, c, seg.get(), sr.get(), dr.get(), e, off.get(), sc.get());
  exclusive_scan(c, off.get(), off.get(), e);
  csize = read_scalar(c, off.get() + e);
  DBuf<uint32_t> out(c, csize);
  HGP_LAUNCH(k_chedge_write, ceil_div(p, 256), 256, 0, c, keys.get(), p, gb, seg.get(), srcf.get(), dstf.get(), sr.get(),
             dr.get(), off.get(), sc.get(), out.get());
  chedges = out.release();
  choff = off.release();
  csrcc = sc.release();
}

void uncoarsen(Ctx& c, const uint32_t* groups, const uint32_t* cpart, uint32_t n, uint32_t* part) {
  HGP_LAUNCH(k_uncoarsen, ceil_div(n, 256), 256, 0, c, groups, cpart, n, part);
}

}  // namespace hgp
