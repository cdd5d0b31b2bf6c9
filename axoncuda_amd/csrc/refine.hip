// refine.hip - parallel FM refinement with size and inbound-set event checks
// (reference refinementRepeats, sources/refinement.cu:18-582; kernels/refinement.cu).
//
// One repeat:
//   1. per-hedge partition lists: for every hedge the distinct partitions of its pins with
//      the count of all pins and of destination pins (replaces the dense E x K
//      pins-per-partition matrix and its sparse bitmap variant);
//   2. gains (fm_refinement_gains_kernel :82-209), exact integers in 2^18 fixed point:
//      gain(q) = saving - W_v + conn(v,q), conn accumulated in an LDS table keyed by
//      partition; non-adjacent targets all share conn = 0 and are ranked by the tie hash;
//   3. ranks: chaining (sources/chaining.cu:19-254) or gain sort (:172-189);
//   4. cascade (:216-322) computed hedge-centrically: each hedge is visited once and every
//      pin's contribution at its own rank is added to seq[rank] (integer atomics), instead
//      of every node re-reading all pins of all its hedges;
//   5. size and inbound-set validity prefixes via radix-sorted events (:239-491);
//   6. best prefix (best_move_functor, headers/data_types.cuh:54-77) and apply (:328-370).
#include <algorithm>
#include <numeric>
#include <optional>
#include <type_traits>

#include "prim.hpp"
#include "stages.hpp"
#include "wave_iter.hpp"

namespace hgp {

namespace {

constexpr uint32_t PP_SMALL = 16;
constexpr uint32_t PP_HIST_CAP = 8192;  // partitions whose inbound counts aggregate in LDS
constexpr uint32_t PP_LDS_CAP = 4096;  // 3 x 16 KiB of dynamic LDS per wave
constexpr uint32_t GAIN_CAP = 1024;
constexpr uint32_t GAIN_LIGHT = 512;
constexpr uint32_t GAIN_BLOCK_MIN = 4096;  // entries above which a node takes a workgroup even for small k
constexpr uint32_t GAIN_TINY = 16;  // partition entries handled by one thread in registers
constexpr uint32_t CAS_SMALL = 16;
constexpr uint32_t CAS_LDS = 2048;
constexpr uint32_t CAS_SORT = 1024;               // pins of a hedge cascaded through sorted events
constexpr uint32_t CAS_SORT_S = 128;              // ... by the small-LDS instance
constexpr uint32_t CAS_EV = 2048;                 // event slots: next power of two >= 2 * CAS_SORT
constexpr uint32_t CAS_SORT_NT = 256;              // threads sorting the events of one such hedge
constexpr uint64_t HEAVY_BATCH_SLOTS = 32ull << 20;

struct Best {
  int32_t ok;
  int32_t sv;
  int64_t seq;
  uint32_t r;
  int32_t iv;
};

__device__ __forceinline__ uint64_t ordered_i64(int64_t g) { return (uint64_t)g ^ 0x8000000000000000ull; }

// ---- 1. per-hedge partition lists ------------------------------------------------------
// Per-hedge partition lists for hedges of at most PP_SMALL pins (one thread per hedge).
// With INB, also the destination counts and the per-partition inbound sizes; those are
// aggregated in an LDS histogram when the partition count fits it (k_lds > 0), so a few
// hot partitions do not serialize millions of global atomics.
template <bool INB>
__global__ void __launch_bounds__(256) k_pp_small(const uint32_t* __restrict__ hedges, const uint64_t* __restrict__ hoff,
                                                  const uint32_t* __restrict__ srcc, const uint32_t* __restrict__ part,
                                                  const float* __restrict__ w, uint32_t E, uint2* __restrict__ pp,
                                                  uint32_t* __restrict__ pp_dcnt, uint32_t* __restrict__ lambda,
                                                  uint4* __restrict__ hinfo, uint32_t* __restrict__ large_flag,
                                                  uint32_t* inbsz, uint32_t k_lds) {
  extern __shared__ uint32_t hist[];
  if (INB && k_lds) {
    for (uint32_t i = threadIdx.x; i < k_lds; i += blockDim.x) hist[i] = 0;
    __syncthreads();
  }
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < E) {
    const uint64_t s = hoff[e], t = hoff[e + 1];
    if (t - s > PP_SMALL) {
      if (large_flag) large_flag[e] = 1u;
    } else {
      if (large_flag) large_flag[e] = 0u;
      const uint32_t len = (uint32_t)(t - s);
      const uint32_t sc = INB ? srcc[e] : 0u;
      // Pins and their partitions are fetched in batches of unconditional loads (indexes
      // clamped to the last pin): two memory round trips per batch instead of two per pin.
      uint32_t pq[PP_SMALL];
#pragma unroll
      for (uint32_t u = 0; u < PP_SMALL; ++u) pq[u] = NONE;
      if (len > 0) {  // an empty hedge has no pin to clamp to
        uint32_t x[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) x[u] = hedges[s + min(u, len - 1)];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) pq[u] = part[x[u]];
      }
      if (len > 4) {
        uint32_t x[PP_SMALL - 4];
#pragma unroll
        for (uint32_t u = 4; u < PP_SMALL; ++u) x[u - 4] = hedges[s + min(u, len - 1)];
#pragma unroll
        for (uint32_t u = 4; u < PP_SMALL; ++u) pq[u] = part[x[u - 4]];
      }
      // (partition, count, destination count) in first-appearance order; static indexes
      // only, so the tables stay in registers
      uint32_t qs[PP_SMALL], cs[PP_SMALL], ds[PP_SMALL];
#pragma unroll
      for (uint32_t k = 0; k < PP_SMALL; ++k) {
        qs[k] = NONE;
        cs[k] = 0;
        ds[k] = 0;
      }
      uint32_t m = 0;
#pragma unroll
      for (uint32_t u = 0; u < PP_SMALL; ++u) {
        if (u < len) {
          const uint32_t q = pq[u];
          const uint32_t d = (INB && u >= sc) ? 1u : 0u;
          bool found = false;
#pragma unroll
          for (uint32_t k = 0; k < PP_SMALL; ++k) {
            const bool hit = qs[k] == q;
            cs[k] += hit ? 1u : 0u;
            if (INB) ds[k] += hit ? d : 0u;
            found |= hit;
          }
          if (!found) {
#pragma unroll
            for (uint32_t k = 0; k < PP_SMALL; ++k) {
              const bool put = k == m;
              qs[k] = put ? q : qs[k];
              cs[k] = put ? 1u : cs[k];
              if (INB) ds[k] = put ? d : ds[k];
            }
            ++m;
          }
        }
      }
#pragma unroll
      for (uint32_t k = 0; k < PP_SMALL; ++k) {
        if (k < m) {
          pp[s + k] = make_uint2(qs[k], cs[k]);
          if (INB) {
            pp_dcnt[s + k] = ds[k];
            if (ds[k]) {
              if (k_lds) atomicAdd(&hist[qs[k]], 1u);
              else atomicAdd(&inbsz[qs[k]], 1u);
            }
          }
        }
      }
      lambda[e] = m;
      // a single entry rides in the record in place of the list start
      const uint64_t x = m == 1 ? (((uint64_t)cs[0] << 32) | qs[0]) : s;
      hinfo[e] = make_uint4((uint32_t)x, (uint32_t)(x >> 32), m, w ? fixed_weight(w[e]) : 0u);
    }
  }
  if (INB && k_lds) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < k_lds; i += blockDim.x)
      if (hist[i]) atomicAdd(&inbsz[i], hist[i]);
  }
}

// Table slots for a hedge of `len` pins spread over at most `k` partitions: a power of two
// of at least twice the distinct partitions (+64), so open addressing never fills.
__host__ __device__ __forceinline__ uint32_t pp_table_cap(uint64_t len, uint32_t k) {
  const uint64_t d = len < k ? len : k;
  uint32_t cap = 128;
  while (cap < 2 * d + 64) cap <<= 1;
  return cap;
}

// One wave per hedge with more than 16 pins.  GMEM = false: the (partition, count, dst
// count) table lives in LDS with `cap` slots; GMEM = true (hedges whose table would not fit
// in LDS): in HBM at gtab + 3 * goff[i], sized by pp_table_cap.  The list is written sorted
// by partition id (the table, compacted into its own first slots and padded to a power of
// two, is bitonic-sorted by the wave), so a partition's entry is found by binary search.
template <bool INB, bool GMEM>
__global__ void __launch_bounds__(64) k_pp_large(const uint32_t* __restrict__ hedges, const uint64_t* __restrict__ hoff,
                                                 const uint32_t* __restrict__ srcc, const uint32_t* __restrict__ part,
                                                 const float* __restrict__ w,
                                                 const uint32_t* __restrict__ list, uint32_t cap, uint32_t k,
                                                 const uint64_t* __restrict__ goff, uint32_t* __restrict__ gtab,
                                                 uint2* __restrict__ pp, uint32_t* __restrict__ pp_dcnt,
                                                 uint32_t* __restrict__ lambda, uint4* __restrict__ hinfo,
                                                 uint32_t* inbsz) {
  extern __shared__ uint32_t lds[];
  const uint32_t e = list[blockIdx.x];
  uint32_t* keys = lds;
  if (GMEM) {
    cap = pp_table_cap(hoff[e + 1] - hoff[e], k);
    keys = gtab + 3 * goff[blockIdx.x];
  }
  uint32_t* cnt = keys + cap;
  uint32_t* dcnt = keys + 2 * cap;
  const uint32_t l = lane_id();
  for (uint32_t i = l; i < cap; i += 64) {
    keys[i] = NONE;
    cnt[i] = 0;
    dcnt[i] = 0;
  }
  __syncthreads();
  const uint64_t s = hoff[e], t = hoff[e + 1];
  const uint32_t sc = srcc[e];
  for (uint64_t i = s + l; i < t; i += 64) {
    const uint32_t q = part[hedges[i]];
    uint32_t h = slot_hash(q) & (cap - 1);
    for (;;) {
      const uint32_t old = atomicCAS(&keys[h], NONE, q);
      if (old == NONE || old == q) break;
      h = (h + 1) & (cap - 1);
    }
    atomicAdd(&cnt[h], 1u);
    if (INB && (i - s) >= sc) atomicAdd(&dcnt[h], 1u);
  }
  __syncthreads();
  // HBM tables were written by atomics at the device cache: read them there too
  auto ld = [](uint32_t* x) { return GMEM ? atomic_load_relaxed(x) : *x; };
  // compaction: the occupied slots, in slot order, into registers of the lanes ahead of the
  // writes (a chunk of 64 slots is read before any of it is overwritten)
  uint32_t base = 0;
  for (uint32_t b = 0; b < cap; b += 64) {
    const uint32_t i = b + l;
    const uint32_t q = ld(&keys[i]);
    const uint32_t c = ld(&cnt[i]), d = ld(&dcnt[i]);
    const bool has = q != NONE;
    const uint64_t m = __ballot(has);
    if (GMEM) __threadfence_block();  // this chunk's reads are done before any lane writes
    __builtin_amdgcn_wave_barrier();
    if (has) {  // base + rank <= b + l: every write lands on a slot this or an earlier chunk read
      const uint32_t o = base + lanes_below(m);
      keys[o] = q;
      cnt[o] = c;
      dcnt[o] = d;
      if (INB && d) atomicAdd(&inbsz[q], 1u);
    }
    if (GMEM) __threadfence_block();
    __builtin_amdgcn_wave_barrier();
    base += (uint32_t)__popcll(m);
  }
  // pad to a power of two and sort the entries by partition id (NONE, the padding, last)
  uint32_t np = 1;
  while (np < base) np <<= 1;
  for (uint32_t i = base + l; i < np; i += 64) keys[i] = NONE;
  if (GMEM) __threadfence_block();
  __builtin_amdgcn_wave_barrier();
  for (uint32_t kk = 2; kk <= np; kk <<= 1) {
    for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
      for (uint32_t tt = l; tt < np / 2; tt += 64) {
        const uint32_t i = (tt / j) * 2 * j + (tt % j), ixj = i + j;
        const uint32_t a = ld(&keys[i]), bq = ld(&keys[ixj]);
        const bool up = (i & kk) == 0;
        if ((a > bq) == up) {
          const uint32_t ca = ld(&cnt[i]), da = ld(&dcnt[i]);
          keys[i] = bq;
          cnt[i] = ld(&cnt[ixj]);
          dcnt[i] = ld(&dcnt[ixj]);
          keys[ixj] = a;
          cnt[ixj] = ca;
          dcnt[ixj] = da;
        }
      }
      if (GMEM) __threadfence_block();
      __builtin_amdgcn_wave_barrier();
    }
  }
  const uint64_t s0 = s;
  for (uint32_t o = l; o < base; o += 64) {
    pp[s0 + o] = make_uint2(ld(&keys[o]), ld(&cnt[o]));
    if (INB) pp_dcnt[s0 + o] = ld(&dcnt[o]);
  }
  if (l == 0) {
    lambda[e] = base;
    // a single entry rides in the record in place of the list start
    const uint64_t x = base == 1 ? (((uint64_t)ld(&cnt[0]) << 32) | ld(&keys[0])) : s0;
    hinfo[e] = make_uint4((uint32_t)x, (uint32_t)(x >> 32), base, w ? fixed_weight(w[e]) : 0u);
  }
}

__global__ void k_flag_longer(const uint32_t* __restrict__ list, uint32_t m, const uint64_t* __restrict__ hoff,
                              uint64_t len, uint32_t* __restrict__ flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t e = list[i];
  flag[i] = hoff[e + 1] - hoff[e] > len ? 1u : 0u;
}
// stable split of list entries: flagged ones to a[], the rest to b[]
__global__ void k_split_list(const uint32_t* __restrict__ flag, const uint64_t* __restrict__ pos,
                             const uint32_t* __restrict__ list, uint32_t m, uint32_t* __restrict__ a,
                             uint32_t* __restrict__ b) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  if (flag[i]) a[pos[i]] = list[i];
  else b[i - pos[i]] = list[i];
}
__global__ void k_pp_gcap(const uint32_t* __restrict__ list, uint32_t m, const uint64_t* __restrict__ hoff, uint32_t k,
                          uint64_t* __restrict__ cap) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t e = list[i];
  cap[i] = pp_table_cap(hoff[e + 1] - hoff[e], k);
}

// ---- 2. gains ---------------------------------------------------------------------------
struct GainArgs {
  const uint32_t* tch;
  const uint64_t* toff;
  const uint4* hinfo;  // per hedge: list start (lo, hi) or, for one entry, the entry; length (lambda); weight
  const uint2* pp;     // per list entry: (partition, pin count)
  const uint32_t* part;
  const uint32_t* sizes;
  const uint32_t* psizes;
  const uint32_t* order;  // partitions by mix32(q + repeat) descending
  uint32_t k;
  uint32_t omega;
  uint32_t discount;
  uint32_t repeat;
  int incc;
  uint32_t* moves;
  int64_t* gains;
  unsigned long long* prof_bytes;  // when profiled: PROF_SLOTS counters of algorithmic bytes
};

__device__ __forceinline__ bool target_ok(const GainArgs& a, uint32_t q, uint32_t p, uint32_t s) {
  return q != p && (uint32_t)(a.psizes[q] + s - s / a.discount) <= a.omega;
}

__device__ __forceinline__ uint32_t table_find(const uint32_t* keys, uint32_t mask, uint32_t q) {
  uint32_t h = slot_hash(q) & mask;
  for (uint32_t probes = 0; probes <= mask; ++probes) {
    const uint32_t k = keys[h];
    if (k == q) return h;
    if (k == NONE) return NONE;
    h = (h + 1) & mask;
  }
  return NONE;
}

// Shared tail of the gain kernels: pick the best target from the (partition -> conn) table.
// Wave-collective; `stride`/`first` give this wave's share of the table slots when several
// waves share one table (heavy path); results combined by the caller.
__device__ __forceinline__ void best_adjacent(const GainArgs& a, const uint32_t* keys, const int64_t* conn,
                                              uint32_t cap, uint32_t first, uint32_t stride, uint32_t p, uint32_t s,
                                              int64_t base, int64_t& bg, uint32_t& bh, uint32_t& bq) {
  bg = INT64_MIN;
  bh = 0;
  bq = NONE;
  for (uint32_t i = first; i < cap; i += stride) {
    const uint32_t q = keys ? keys[i] : i;  // keys == nullptr: table indexed by partition
    if (q == NONE) continue;
    const int64_t c = conn[i];
    if (c <= 0 || !target_ok(a, q, p, s)) continue;
    const int64_t g = base + c;
    const uint32_t h = mix32(q + a.repeat);
    if (bq == NONE || g > bg || (g == bg && h > bh)) {
      bg = g;
      bh = h;
      bq = q;
    }
  }
}

__device__ __forceinline__ void wave_best(int64_t& bg, uint32_t& bh, uint32_t& bq) {
  // reduce (gain, hash) lexicographically across the wave; lanes without a target carry bq = NONE
  int64_t g = bq == NONE ? INT64_MIN : bg;
  const int64_t gm = wave_max(g);
  uint32_t h = (bq != NONE && g == gm) ? bh : 0u;
  const bool eligible = bq != NONE && g == gm;
  const uint64_t hk = eligible ? (((uint64_t)h << 32) | 1u) : 0ull;  // (hash, present)
  const uint64_t hm = wave_max(hk);
  const bool mine = eligible && hk == hm;
  const uint64_t who = __ballot(mine);
  if (who == 0) {
    bq = NONE;
    return;
  }
  const int src = __ffsll((long long)who) - 1;
  bq = __shfl(bq, src, 64);
  bh = __shfl(bh, src, 64);
  bg = gm;
}

// non-adjacent fallback: first partition in hash order that is a valid target and has no
// positive connection to v (all such targets share the gain `base`)
__device__ __forceinline__ uint32_t first_nonadjacent(const GainArgs& a, const uint32_t* keys, const int64_t* conn,
                                                      uint32_t mask, uint32_t p, uint32_t s) {
  const uint32_t l = lane_id();
  for (uint32_t b = 0; b < a.k; b += 64) {
    const uint32_t i = b + l;
    bool ok = false;
    uint32_t q = NONE;
    if (i < a.k) {
      q = a.order[i];
      if (target_ok(a, q, p, s)) {
        const uint32_t slot = keys ? table_find(keys, mask, q) : q;
        ok = slot == NONE || conn[slot] <= 0;
      }
    }
    const uint64_t m = __ballot(ok);
    if (m) return __shfl(q, __ffsll((long long)m) - 1, 64);
  }
  return NONE;
}

// Nodes whose touching hedges list at most GAIN_TINY partition entries: one thread per node,
// the (partition, connection) table held in registers.
__device__ __forceinline__ void gains_tiny_node(const GainArgs& a, uint32_t v) {
  const uint32_t p = a.part[v], s = a.sizes[v];
  uint32_t qs[GAIN_TINY];
  int64_t cs[GAIN_TINY];
#pragma unroll
  for (uint32_t k = 0; k < GAIN_TINY; ++k) {
    qs[k] = NONE;
    cs[k] = 0;
  }
  uint32_t m = 0;
  int64_t saving = 0, Wv = 0;
  for (uint64_t j = a.toff[v]; j < a.toff[v + 1]; ++j) {
    const uint4 hi = a.hinfo[a.tch[j]];
    const int64_t We = (int64_t)hi.w;
    Wv += We;
    const uint64_t b = ((uint64_t)hi.y << 32) | hi.x;
    const uint32_t lam = hi.z;
    for (uint32_t x = 0; x < lam; ++x) {
      const uint2 pe = lam == 1 ? make_uint2(hi.x, hi.y) : a.pp[b + x];  // one entry: in the record
      const uint32_t q = pe.x;
      if (q == p) {
        const uint32_t c = pe.y;
        if (a.incc) saving += (int64_t)((uint64_t)We / ((uint64_t)c * c));
        else if (c == 1) saving += We;
        continue;
      }
      bool found = false;
#pragma unroll
      for (uint32_t k = 0; k < GAIN_TINY; ++k) {
        const bool hit = qs[k] == q;
        cs[k] += hit ? We : 0;
        found |= hit;
      }
      if (!found) {
#pragma unroll
        for (uint32_t k = 0; k < GAIN_TINY; ++k) {
          const bool put = k == m;
          qs[k] = put ? q : qs[k];
          cs[k] = put ? We : cs[k];
        }
        ++m;
      }
    }
  }
  const int64_t base = saving - Wv;
  int64_t bg = INT64_MIN;
  uint32_t bh = 0, bq = NONE;
#pragma unroll
  for (uint32_t k = 0; k < GAIN_TINY; ++k) {
    const uint32_t q = qs[k];
    if (q == NONE || cs[k] <= 0 || !target_ok(a, q, p, s)) continue;
    const int64_t g = base + cs[k];
    const uint32_t h = mix32(q + a.repeat);
    if (bq == NONE || g > bg || (g == bg && h > bh)) {
      bg = g;
      bh = h;
      bq = q;
    }
  }
  if (bq == NONE) {  // first valid partition in hash order without a positive connection
    for (uint32_t x = 0; x < a.k; ++x) {
      const uint32_t q = a.order[x];
      if (!target_ok(a, q, p, s)) continue;
      bool adjacent = false;
#pragma unroll
      for (uint32_t k = 0; k < GAIN_TINY; ++k) adjacent |= qs[k] == q && cs[k] > 0;
      if (!adjacent) {
        bq = q;
        break;
      }
    }
    bg = base;
  }
  a.moves[v] = bq;
  a.gains[v] = bq == NONE ? 0 : bg;
}

// One wave per node.  With at most GAIN_CAP partitions the LDS table is indexed by partition
// (k slots); otherwise it is an open-addressing table sized to twice the node's possible
// distinct partitions.
// Dynamic LDS: 4 waves x capmax slots of (int64 connection, uint32 key).  Returns this lane's
// share of the node's algorithmic bytes: 36 (node record, offsets, result) + 20 per touching
// hedge (list entry and hedge record) + 8 per partition entry.
__device__ __forceinline__ uint64_t gains_light_node(const GainArgs& a, uint32_t v, const uint64_t* __restrict__ bound,
                                                 uint32_t capmax, int64_t* s_conn_all, uint32_t wv, uint32_t l) {
  const bool direct = a.k <= GAIN_CAP;
  uint32_t cap = a.k;
  if (!direct) {
    const uint64_t d = bound[v] < (uint64_t)a.k ? bound[v] : (uint64_t)a.k;
    cap = 64;
    while (cap < 2 * d + 64 && cap < GAIN_CAP) cap <<= 1;
  }
  uint32_t* keys = direct ? nullptr : (uint32_t*)(s_conn_all + 4 * capmax) + wv * capmax;
  int64_t* conn = s_conn_all + wv * capmax;
  for (uint32_t i = l; i < cap; i += 64) {
    if (!direct) keys[i] = NONE;
    conn[i] = 0;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const uint32_t p = a.part[v], s = a.sizes[v];
  const uint64_t t0 = a.toff[v], t1 = a.toff[v + 1];
  int64_t saving = 0, Wv = 0;
  uint64_t entries = 0;
  Wv += wave_visit_hedge_lists(a.tch, t0, t1, a.hinfo, a.pp, [&](const EntryVisit& ev) {
        if (!ev.act) return;
        ++entries;
        const uint2 pe = ev.pe;
        const uint32_t q = pe.x;
        const int64_t We = (int64_t)ev.w;
        if (q == p) {
          const uint32_t c = pe.y;
          if (a.incc) saving += (int64_t)((uint64_t)We / ((uint64_t)c * c));
          else if (c == 1) saving += We;
          return;
        }
        uint32_t h = q;
        if (!direct) {
          h = slot_hash(q) & (cap - 1);
          for (;;) {
            const uint32_t old = atomicCAS(&keys[h], NONE, q);
            if (old == NONE || old == q) break;
            h = (h + 1) & (cap - 1);
          }
        }
        atomicAdd((unsigned long long*)&conn[h], (unsigned long long)We);
      });
  saving = wave_sum(saving);
  Wv = wave_sum(Wv);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const int64_t base = saving - Wv;
  int64_t bg;
  uint32_t bh, bq;
  best_adjacent(a, keys, conn, cap, l, 64, p, s, base, bg, bh, bq);
  wave_best(bg, bh, bq);
  if (bq == NONE) {
    bq = first_nonadjacent(a, keys, conn, cap - 1, p, s);
    bg = base;
  }
  if (l == 0) {
    a.moves[v] = bq;
    a.gains[v] = bq == NONE ? 0 : bg;
  }
  return 8 * entries + (l == 0 ? 36 + 20 * (t1 - t0) : 0);
}

// Persistent: the grid strides over the `*count` listed nodes (the count stays on the device).
__global__ void __launch_bounds__(256) k_gains_light(GainArgs a, const uint32_t* __restrict__ list,
                                                     const uint32_t* __restrict__ count,
                                                     const uint64_t* __restrict__ bound, uint32_t capmax) {
  extern __shared__ int64_t s_conn_all[];
  const uint32_t wv = threadIdx.x / 64, l = lane_id();
  const uint32_t nlist = *count;
  uint64_t bytes = 0;
  for (uint32_t idx = blockIdx.x * 4 + wv; idx < nlist; idx += gridDim.x * 4) {
    bytes += gains_light_node(a, list[idx], bound, capmax, s_conn_all, wv, l);
    // every lane has read the table before the next node clears it
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  if (a.prof_bytes) {
    bytes = wave_sum(bytes);
    if (l == 0 && bytes) atomicAdd(&a.prof_bytes[blockIdx.x % PROF_SLOTS], (unsigned long long)bytes);
  }
}

__global__ void __launch_bounds__(256) k_gains_heavy(GainArgs a, const uint32_t* __restrict__ list,
                                                     const uint64_t* __restrict__ tab_off, uint32_t* __restrict__ gkeys,
                                                     int64_t* __restrict__ gconn) {
  __shared__ int64_t s_sum[2][4];
  __shared__ int64_t s_bg[4];
  __shared__ uint32_t s_bh[4], s_bq[4];
  const uint32_t v = list[blockIdx.x];
  const uint64_t tb = tab_off[blockIdx.x];
  const uint32_t cap = (uint32_t)(tab_off[blockIdx.x + 1] - tb);
  uint32_t* keys = gkeys + tb;
  int64_t* conn = gconn + tb;
  for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x) {
    keys[i] = NONE;
    conn[i] = 0;
  }
  __threadfence_block();
  __syncthreads();
  const uint32_t wv = threadIdx.x / 64, l = lane_id();
  const uint32_t p = a.part[v], s = a.sizes[v];
  const uint64_t t0 = a.toff[v], t1 = a.toff[v + 1];
  const uint64_t cnt = t1 - t0;
  const uint64_t lo = t0 + cnt * wv / 4, hi = t0 + cnt * (wv + 1) / 4;
  int64_t saving = 0, Wv = 0;
  Wv += wave_visit_hedge_lists(a.tch, lo, hi, a.hinfo, a.pp, [&](const EntryVisit& ev) {
        if (!ev.act) return;
        const uint2 pe = ev.pe;
        const uint32_t q = pe.x;
        const int64_t We = (int64_t)ev.w;
        if (q == p) {
          const uint32_t c = pe.y;
          if (a.incc) saving += (int64_t)((uint64_t)We / ((uint64_t)c * c));
          else if (c == 1) saving += We;
          return;
        }
        uint32_t h = slot_hash(q) & (cap - 1);
        for (;;) {
          const uint32_t old = atomicCAS(&keys[h], NONE, q);
          if (old == NONE || old == q) break;
          h = (h + 1) & (cap - 1);
        }
        atomicAdd((unsigned long long*)&conn[h], (unsigned long long)We);
      });
  saving = wave_sum(saving);
  Wv = wave_sum(Wv);
  if (l == 0) {
    s_sum[0][wv] = saving;
    s_sum[1][wv] = Wv;
  }
  __threadfence();
  __syncthreads();
  const int64_t base = (s_sum[0][0] + s_sum[0][1] + s_sum[0][2] + s_sum[0][3]) -
                       (s_sum[1][0] + s_sum[1][1] + s_sum[1][2] + s_sum[1][3]);
  int64_t bg;
  uint32_t bh, bq;
  best_adjacent(a, keys, conn, cap, threadIdx.x, blockDim.x, p, s, base, bg, bh, bq);
  wave_best(bg, bh, bq);
  if (l == 0) {
    s_bg[wv] = bg;
    s_bh[wv] = bh;
    s_bq[wv] = bq;
  }
  __syncthreads();
  if (wv == 0) {
    int64_t g = INT64_MIN;
    uint32_t h = 0, q = NONE;
    for (int i = 0; i < 4; ++i) {
      if (s_bq[i] == NONE) continue;
      if (q == NONE || s_bg[i] > g || (s_bg[i] == g && s_bh[i] > h)) {
        g = s_bg[i];
        h = s_bh[i];
        q = s_bq[i];
      }
    }
    if (q == NONE) {
      q = first_nonadjacent(a, keys, conn, cap - 1, p, s);
      g = base;
    }
    if (l == 0) {
      a.moves[v] = q;
      a.gains[v] = q == NONE ? 0 : g;
    }
  }
}

// Nodes above the light tier when k > GAIN_LIGHT: one 1024-thread workgroup per node with a
// (partition -> conn) table filling the LDS.  DIRECT (k <= GB_DIRECT): indexed by partition.
// Hashed: GB_SLOTS slots (not a power of two: multiply-high index).  A new partition is
// inserted only while fewer than GB_LIMIT are present; a node that meets one more sets the
// overflow flag and is listed in `ovf` for the HBM-table kernel.  At most one insert per
// thread can pass the check concurrently, so GB_LIMIT + 1024 < GB_SLOTS keeps a free slot
// for every probe.  The slots a node inserts are also listed (GB_LIMIT + 1024 entries), so it
// ranks and clears only those instead of the whole table.
constexpr uint32_t LIGHT_GRID = 16384;  // workgroups of the persistent light kernel (4 nodes each at a time)
constexpr uint32_t BLOCK_GRID = 2048;   // workgroups of the persistent block kernel
constexpr uint64_t BOUND_UNKNOWN = ~0ull;  // a node routed by degree: its bound was not summed
constexpr uint32_t WIDE_LIGHT_DEG = 256;   // ... and it takes a wave (else a workgroup) up to this degree
constexpr uint32_t GB_DIRECT = 20160;  // 8 bytes per slot
constexpr uint32_t GB_SLOTS = 10496;   // 12 bytes per slot, and 3 for the list of inserted slots
constexpr uint32_t GB_LIMIT = GB_SLOTS / 4 * 3;
static_assert(GB_LIMIT + 1024 < GB_SLOTS, "the table must keep a free slot for every in-flight insert");

__device__ __forceinline__ uint32_t gb_find(const uint32_t* keys, uint32_t q) {
  uint32_t h = __umulhi(slot_hash(q), GB_SLOTS);
  for (;;) {
    const uint32_t k = keys[h];
    if (k == q) return h;
    if (k == NONE) return NONE;
    h = h + 1 == GB_SLOTS ? 0u : h + 1;
  }
}

template <bool DIRECT>
__global__ void __launch_bounds__(1024) k_gains_block(GainArgs a, const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ count, uint32_t* __restrict__ ovf,
                                                      uint32_t* __restrict__ novf) {
  constexpr uint32_t SLOTS = DIRECT ? GB_DIRECT : GB_SLOTS;
  __shared__ int64_t s_conn[SLOTS];
  __shared__ uint32_t s_keys[DIRECT ? 1 : SLOTS];
  __shared__ uint32_t s_ins[DIRECT ? 1 : GB_LIMIT + 1024];  // slots inserted for this node
  __shared__ int64_t s_sum[2][16];
  __shared__ int64_t s_bg[16];
  __shared__ uint32_t s_bh[16], s_bq[16];
  __shared__ uint32_t s_state[2];  // partitions inserted, overflow flag
  __shared__ uint32_t s_first;
  const uint32_t cap = DIRECT ? a.k : SLOTS;
  const uint32_t nlist = *count;  // persistent: the grid strides over the listed nodes
  if (!DIRECT) {  // the hashed table is cleared once; each node then clears the slots it used
    for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x) {
      s_conn[i] = 0;
      s_keys[i] = NONE;
    }
  }
  // the hashed table's slots used by the node: cleared (with the node's state reset after a barrier)
  auto clear_used = [&] {
    __syncthreads();  // every thread is done with the table and s_state
    const uint32_t used = min(s_state[0], GB_LIMIT + 1024);
    for (uint32_t i = threadIdx.x; i < used; i += blockDim.x) {
      const uint32_t h = s_ins[i];
      s_keys[h] = NONE;
      s_conn[h] = 0;
    }
  };
  for (uint32_t bi = blockIdx.x; bi < nlist; bi += gridDim.x) {
  const uint32_t v = list[bi];
  if (DIRECT)
    for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x) s_conn[i] = 0;
  __syncthreads();  // the previous node's clearing is done before its state is reset
  if (threadIdx.x == 0) {
    s_state[0] = 0;
    s_state[1] = 0;
    s_first = UINT32_MAX;
  }
  __syncthreads();
  const uint32_t wv = threadIdx.x / 64, l = lane_id();
  const uint32_t p = a.part[v], s = a.sizes[v];
  const uint64_t t0 = a.toff[v], t1 = a.toff[v + 1];
  const uint64_t cnt = t1 - t0;
  const uint64_t lo = t0 + cnt * wv / 16, hi = t0 + cnt * (wv + 1) / 16;
  int64_t saving = 0, Wv = 0;
  Wv += wave_visit_hedge_lists(a.tch, lo, hi, a.hinfo, a.pp, [&](const EntryVisit& ev) {
    if (!ev.act) return;
    const uint2 pe = ev.pe;
    const uint32_t q = pe.x;
    const int64_t We = (int64_t)ev.w;
    if (q == p) {
      const uint32_t c = pe.y;
      if (a.incc) saving += (int64_t)((uint64_t)We / ((uint64_t)c * c));
      else if (c == 1) saving += We;
      return;
    }
    uint32_t h = q;
    if (!DIRECT) {
      h = __umulhi(slot_hash(q), GB_SLOTS);
      for (;;) {
        const uint32_t k = atomic_load_relaxed(&s_keys[h]);
        if (k == q) break;
        if (k == NONE) {
          if (atomic_load_relaxed(&s_state[0]) >= GB_LIMIT) {
            atomicOr(&s_state[1], 1u);  // table full enough: the node goes to the HBM kernel
            return;
          }
          const uint32_t old = atomicCAS(&s_keys[h], NONE, q);
          if (old == NONE) {
            s_ins[atomicAdd(&s_state[0], 1u)] = h;  // below GB_LIMIT + 1024: one in flight per thread
            break;
          }
          if (old == q) break;
        }
        h = h + 1 == GB_SLOTS ? 0u : h + 1;
      }
    }
    atomicAdd((unsigned long long*)&s_conn[h], (unsigned long long)We);
  });
  saving = wave_sum(saving);
  Wv = wave_sum(Wv);
  if (l == 0) {
    s_sum[0][wv] = saving;
    s_sum[1][wv] = Wv;
  }
  __syncthreads();
  const bool overflow = !DIRECT && s_state[1];
  if (overflow) {
    if (threadIdx.x == 0) ovf[atomicAdd(novf, 1u)] = v;
    clear_used();
    continue;
  }
  int64_t base = 0;
  for (int i = 0; i < 16; ++i) base += s_sum[0][i] - s_sum[1][i];
  int64_t bg;
  uint32_t bh, bq;
  if (DIRECT) {
    best_adjacent(a, nullptr, s_conn, cap, threadIdx.x, blockDim.x, p, s, base, bg, bh, bq);
  } else {  // the inserted slots only
    bg = INT64_MIN;
    bh = 0;
    bq = NONE;
    const uint32_t used = s_state[0];
    for (uint32_t i = threadIdx.x; i < used; i += blockDim.x) {
      const uint32_t hs = s_ins[i];
      const uint32_t q = s_keys[hs];
      const int64_t c = s_conn[hs];
      if (c <= 0 || !target_ok(a, q, p, s)) continue;
      const int64_t g = base + c;
      const uint32_t h = mix32(q + a.repeat);
      if (bq == NONE || g > bg || (g == bg && h > bh)) {
        bg = g;
        bh = h;
        bq = q;
      }
    }
  }
  wave_best(bg, bh, bq);
  if (l == 0) {
    s_bg[wv] = bg;
    s_bh[wv] = bh;
    s_bq[wv] = bq;
  }
  __syncthreads();
  int64_t g = INT64_MIN;
  uint32_t h = 0, q = NONE;
  for (int i = 0; i < 16; ++i) {
    if (s_bq[i] == NONE) continue;
    if (q == NONE || s_bg[i] > g || (s_bg[i] == g && s_bh[i] > h)) {
      g = s_bg[i];
      h = s_bh[i];
      q = s_bq[i];
    }
  }
  if (q == NONE) {
    // non-adjacent fallback: the first valid target in hash order without a positive
    // connection, searched by the whole workgroup 1024 partitions at a time
    for (uint32_t b = 0; b < a.k; b += blockDim.x) {
      const uint32_t i = b + threadIdx.x;
      if (i < a.k) {
        const uint32_t t = a.order[i];
        if (target_ok(a, t, p, s)) {
          const uint32_t slot = DIRECT ? t : gb_find(s_keys, t);
          if (slot == NONE || s_conn[slot] <= 0) atomicMin(&s_first, i);
        }
      }
      __syncthreads();
      const bool found = s_first != UINT32_MAX;
      __syncthreads();  // every thread has read s_first before the next round may lower it
      if (found) break;
    }
    if (s_first != UINT32_MAX) q = a.order[s_first];
    g = base;
  }
  if (threadIdx.x == 0) {
    a.moves[v] = q;
    a.gains[v] = q == NONE ? 0 : g;
  }
  if (!DIRECT) clear_used();
  }
}

// Appends the lanes whose flag is set to a device list: one atomic per workgroup and list,
// then wave and lane offsets.  Every thread of the 256-thread workgroup must call it.
template <int NL>
__device__ __forceinline__ void append_lists(const bool (&to)[NL], uint32_t v, uint32_t* const (&lists)[NL],
                                             uint32_t* __restrict__ counts) {
  __shared__ uint32_t s_cnt[NL][4], s_base[NL];
  const uint32_t l = lane_id(), wv = threadIdx.x / 64;
  uint64_t m[NL];
#pragma unroll
  for (int x = 0; x < NL; ++x) {
    m[x] = __ballot(to[x]);
    if (l == 0) s_cnt[x][wv] = (uint32_t)__popcll(m[x]);
  }
  __syncthreads();
  if (threadIdx.x < NL) {
    const uint32_t t = s_cnt[threadIdx.x][0] + s_cnt[threadIdx.x][1] + s_cnt[threadIdx.x][2] + s_cnt[threadIdx.x][3];
    s_base[threadIdx.x] = t ? atomicAdd(&counts[threadIdx.x], t) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int x = 0; x < NL; ++x) {
    uint32_t base = s_base[x];
    for (uint32_t w = 0; w < wv; ++w) base += s_cnt[x][w];
    if (to[x]) lists[x][base + lanes_below(m[x])] = v;
  }
  __syncthreads();  // the shared counts are reused by the next call
}

// Routes every node by bound = partition entries listed by its touching hedges: bound <=
// GAIN_TINY is scored right here (thread per node, registers); the others are appended to
// the light list (wave per node; up to GAIN_BLOCK_MIN entries when k <= GAIN_LIGHT) or the
// block list (workgroup per node), with the counts left on the device.  When k <= GAIN_LIGHT
// the light tables are indexed by partition and the bound only balances the work, so a node
// touching more than GAIN_TINY hedges (it cannot be tiny) is routed by its degree instead of
// being summed by one thread, and its bound is left unknown (BOUND_UNKNOWN).  List order
// varies from run to run; each node's result does not depend on it nor on the route.
__global__ void __launch_bounds__(256) k_gains_route(GainArgs a, uint32_t n, uint64_t* __restrict__ bound,
                                                     uint32_t* __restrict__ light, uint32_t* __restrict__ block,
                                                     uint32_t* __restrict__ counts) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t t0 = 0, t1 = 0;
  if (v < n) {
    t0 = a.toff[v];
    t1 = a.toff[v + 1];
  }
  const uint64_t deg = t1 - t0;
  const bool by_degree = v < n && a.k <= GAIN_LIGHT && deg > GAIN_TINY;
  uint64_t b = 0;
  if (v < n && !by_degree && deg > GAIN_TINY) {
    for (uint64_t j = t0; j < t1; ++j) b += a.hinfo[a.tch[j]].z;
  } else if (v < n && !by_degree && deg > 0) {
    // at most GAIN_TINY touching hedges: unconditional loads (indexes clamped to the last
    // one), 4 and then the rest, so the sum costs two memory round trips per batch
    const uint32_t d = (uint32_t)deg;
    uint32_t e[GAIN_TINY];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) e[u] = a.tch[t0 + min(u, d - 1)];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t z = a.hinfo[e[u]].z;
      b += u < d ? z : 0u;
    }
    if (d > 4) {
#pragma unroll
      for (uint32_t u = 4; u < GAIN_TINY; ++u) e[u] = a.tch[t0 + min(u, d - 1)];
#pragma unroll
      for (uint32_t u = 4; u < GAIN_TINY; ++u) {
        const uint32_t z = a.hinfo[e[u]].z;
        b += u < d ? z : 0u;
      }
    }
  }
  const bool tiny = v < n && !by_degree && b <= GAIN_TINY;
  if (tiny) gains_tiny_node(a, v);
  const bool rest = v < n && !tiny;
  if (rest) bound[v] = by_degree ? BOUND_UNKNOWN : b;
  const bool to_light =
      rest && (by_degree ? deg <= WIDE_LIGHT_DEG : (b <= GAIN_LIGHT || (a.k <= GAIN_LIGHT && b <= GAIN_BLOCK_MIN)));
  const bool to_block = rest && !to_light;
  // list slots: one atomic per workgroup and list, then wave and lane offsets
  __shared__ uint32_t s_cnt[2][4], s_base[2];
  const uint32_t l = lane_id(), wv = threadIdx.x / 64;
  const uint64_t ml = __ballot(to_light), mb = __ballot(to_block);
  if (l == 0) {
    s_cnt[0][wv] = (uint32_t)__popcll(ml);
    s_cnt[1][wv] = (uint32_t)__popcll(mb);
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const uint32_t t = s_cnt[threadIdx.x][0] + s_cnt[threadIdx.x][1] + s_cnt[threadIdx.x][2] + s_cnt[threadIdx.x][3];
    s_base[threadIdx.x] = t ? atomicAdd(&counts[threadIdx.x], t) : 0u;
  }
  __syncthreads();
  uint32_t base_l = s_base[0], base_b = s_base[1];
  for (uint32_t w = 0; w < wv; ++w) {
    base_l += s_cnt[0][w];
    base_b += s_cnt[1][w];
  }
  if (to_light) light[base_l + lanes_below(ml)] = v;
  if (to_block) block[base_b + lanes_below(mb)] = v;
}

// ---- k-way gains from partition masks (k <= MASK_K) ------------------------------------
// With at most 64 partitions a hedge's partition list fits in two 64-bit masks: the
// partitions among its pins and those holding exactly one of them.  For a node v in p:
// Wv = sum of its hedge weights, saving = sum of the weights whose single-pin mask has p
// (v is then the hedge's only pin in p), conn(q) = sum of the weights whose mask has q; the
// gain of moving to q is saving - Wv + conn(q), exactly what the list kernels compute, at one
// 16-byte record per touching hedge whatever the hedge's connectivity.
constexpr uint32_t MASK_K = 64;
constexpr uint32_t MASK_SMALL_K = 8;  // partitions up to which the wide kernels sum in registers

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int o) {
  const uint32_t lo = __shfl_xor((uint32_t)v, o, 64), hi = __shfl_xor((uint32_t)(v >> 32), o, 64);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ void mask_add(uint64_t& s1, uint64_t& s2, uint32_t q) {
  const uint64_t bit = 1ull << q;
  s2 |= s1 & bit;  // seen before: at least two pins
  s1 |= bit;
}
// Hedges of at most PP_SMALL pins, one thread each (batched unconditional loads, indexes
// clamped to the last pin); longer ones are flagged for k_hmask_large.
__global__ void __launch_bounds__(256) k_hmask_small(const uint32_t* __restrict__ hedges,
                                                     const uint64_t* __restrict__ hoff,
                                                     const uint32_t* __restrict__ part, uint32_t E,
                                                     uint4* __restrict__ hm, uint32_t* __restrict__ large_flag) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const uint64_t s = hoff[e], t = hoff[e + 1];
  const uint64_t hl = t - s;
  if (large_flag) large_flag[e] = hl > PP_SMALL ? 1u : 0u;
  if (hl > PP_SMALL) return;
  const uint32_t len = (uint32_t)hl;
  uint64_t s1 = 0, s2 = 0;
  if (len > 0) {
    uint32_t x[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) x[u] = hedges[s + min(u, len - 1)];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t q = part[x[u]];
      if (u < len) mask_add(s1, s2, q);
    }
    if (len > 4) {
      uint32_t y[PP_SMALL - 4];
#pragma unroll
      for (uint32_t u = 4; u < PP_SMALL; ++u) y[u - 4] = hedges[s + min(u, len - 1)];
#pragma unroll
      for (uint32_t u = 4; u < PP_SMALL; ++u) {
        const uint32_t q = part[y[u - 4]];
        if (u < len) mask_add(s1, s2, q);
      }
    }
  }
  const uint64_t single = s1 & ~s2;
  hm[e] = make_uint4((uint32_t)s1, (uint32_t)(s1 >> 32), (uint32_t)single, (uint32_t)(single >> 32));
}
// One wave per listed hedge: lanes stride over the pins, then the (seen, seen twice) masks
// are combined across lanes.
__global__ void __launch_bounds__(64) k_hmask_large(const uint32_t* __restrict__ hedges,
                                                    const uint64_t* __restrict__ hoff,
                                                    const uint32_t* __restrict__ part,
                                                    const uint32_t* __restrict__ list, uint4* __restrict__ hm) {
  const uint32_t e = list[blockIdx.x];
  const uint32_t l = lane_id();
  const uint64_t s = hoff[e], t = hoff[e + 1];
  uint64_t s1 = 0, s2 = 0;
  for (uint64_t i = s + l; i < t; i += 64) mask_add(s1, s2, part[hedges[i]]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t o1 = shfl_xor_u64(s1, o), o2 = shfl_xor_u64(s2, o);
    s2 |= o2 | (s1 & o1);
    s1 |= o1;
  }
  if (l == 0) {
    const uint64_t single = s1 & ~s2;
    hm[e] = make_uint4((uint32_t)s1, (uint32_t)(s1 >> 32), (uint32_t)single, (uint32_t)(single >> 32));
  }
}

// The mask path's node list is split by XCD: the router's workgroups take node ranges through
// xcd_block, and the nodes one XCD routes go to that XCD's own sub-list, starting at the
// first node of its range (a range holds at most that many listed nodes) with a counter of its
// own.  The wave kernels' workgroups on XCD x consume sub-list x, so the hedge records of
// neighbouring nodes are fetched into one L2.
// Below XCD_SPLIT_GRID router workgroups the split would leave XCDs without nodes, so there is
// one list (number 0) that every workgroup consumes.
constexpr uint32_t XCDS = 8, XCD_SPLIT_GRID = 64;
__device__ __forceinline__ uint32_t xcd_list_base(uint32_t x, uint32_t route_grid) {
  const uint32_t q = route_grid / XCDS, r = route_grid % XCDS;
  return (x * q + (x < r ? x : r)) * 256u;  // first node of XCD x's range (256 per workgroup)
}
// (list, first workgroup index, workgroups) a mask-kernel workgroup consumes
struct SubList {
  const uint32_t* list;
  uint32_t count, first, step;
};
__device__ __forceinline__ SubList sub_list(const uint32_t* list, const uint32_t* count, uint32_t route_grid) {
  if (route_grid < XCD_SPLIT_GRID) return SubList{list, count[0], blockIdx.x, gridDim.x};
  const uint32_t x = blockIdx.x % XCDS;
  return SubList{list + xcd_list_base(x, route_grid), count[x], blockIdx.x / XCDS, (gridDim.x - x + XCDS - 1) / XCDS};
}

// Routing for the mask path: a node touching at most GAIN_TINY hedges is scored right here,
// one thread per node with its hedges' masks and weights in registers (batched loads): the
// candidate targets are the bits of the masks' union, each scored by one pass over the
// masks.  Every other node goes to the wave kernel's list.
__global__ void __launch_bounds__(256) k_gains_route_mask(GainArgs a, const uint4* __restrict__ hm,
                                                          const float* __restrict__ w, uint32_t n,
                                                          uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
  // partition sizes and the hash order, for the candidates' room checks and the fallback
  __shared__ uint32_t s_ps[MASK_K], s_ord[MASK_K];
  if (threadIdx.x < a.k) {
    s_ps[threadIdx.x] = a.psizes[threadIdx.x];
    s_ord[threadIdx.x] = a.order[threadIdx.x];
  }
  __syncthreads();
  const bool split = gridDim.x >= XCD_SPLIT_GRID;
  const uint32_t v = (split ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t t0 = 0, t1 = 0;
  if (v < n) {
    t0 = a.toff[v];
    t1 = a.toff[v + 1];
  }
  const uint64_t deg = t1 - t0;
  const bool small = v < n && deg <= GAIN_TINY;
  if (small) {
    const uint32_t p = a.part[v], s = a.sizes[v];
    const uint32_t d = (uint32_t)deg;
    uint64_t mk[GAIN_TINY];
    uint32_t wt[GAIN_TINY];
    int64_t saving = 0, Wv = 0;
#pragma unroll
    for (uint32_t u = 0; u < GAIN_TINY; ++u) {
      mk[u] = 0;
      wt[u] = 0;
    }
    // records of hedges u0 .. u0 + B - 1: all loads issued first, then consumed
    auto batch = [&](auto B_, uint32_t u0) {
      constexpr uint32_t B = decltype(B_)::value;
      uint32_t e[B];
#pragma unroll
      for (uint32_t u = 0; u < B; ++u) e[u] = a.tch[t0 + min(u0 + u, d - 1)];
      uint4 r[B];
      float wf[B];
#pragma unroll
      for (uint32_t u = 0; u < B; ++u) {
        r[u] = hm[e[u]];
        wf[u] = w[e[u]];
      }
#pragma unroll
      for (uint32_t u = 0; u < B; ++u) {
        const bool in = u0 + u < d;
        const uint32_t we = in ? fixed_weight(wf[u]) : 0u;
        const uint64_t m = in ? ((uint64_t)r[u].y << 32) | r[u].x : 0ull;
        const uint64_t sg = ((uint64_t)r[u].w << 32) | r[u].z;
        mk[u0 + u] = m;
        wt[u0 + u] = we;
        Wv += we;
        saving += ((sg >> p) & 1ull) ? (int64_t)we : 0;  // v is the hedge's only pin in p
      }
    };
    if (d > 0) batch(std::integral_constant<uint32_t, 4>{}, 0);
    if (d > 4) batch(std::integral_constant<uint32_t, GAIN_TINY - 4>{}, 4);
    // entries past a node's degree are zero: the sums stop at the largest degree among the wave's
    // tiny nodes, rounded up to 4 (uniform over the active lanes)
    const uint32_t dmax = __ballot(d > 12) ? 16u : __ballot(d > 8) ? 12u : __ballot(d > 4) ? 8u : 4u;
    auto conn_of = [&](uint32_t q) {
      int64_t c = 0;
#pragma unroll
      for (uint32_t g = 0; g < GAIN_TINY / 4; ++g) {
        if (4 * g >= dmax) break;
#pragma unroll
        for (uint32_t u = 4 * g; u < 4 * g + 4; ++u) c += ((mk[u] >> q) & 1ull) ? (int64_t)wt[u] : 0;
      }
      return c;
    };
    const int64_t base = saving - Wv;
    auto fits = [&](uint32_t q) { return q != p && (uint32_t)(s_ps[q] + s - s / a.discount) <= a.omega; };
    // partitions with a positive connection: in a mask of positive weight (weights are >= 0)
    uint64_t pos = 0;
#pragma unroll
    for (uint32_t u = 0; u < GAIN_TINY; ++u) pos |= wt[u] ? mk[u] : 0ull;
    pos &= ~(1ull << p);
    int64_t bg = INT64_MIN;
    uint32_t bh = 0, bq = NONE;
    for (uint64_t cand = pos; cand;) {  // ascending; (gain, hash) decides, the hash is a bijection
      const uint32_t q = (uint32_t)__ffsll((long long)cand) - 1;
      cand &= cand - 1;
      if (!fits(q)) continue;
      const int64_t c = conn_of(q);
      const int64_t g = base + c;
      const uint32_t h = mix32(q + a.repeat);
      if (bq == NONE || g > bg || (g == bg && h > bh)) {
        bg = g;
        bh = h;
        bq = q;
      }
    }
    if (bq == NONE) {  // first valid partition in hash order without a positive connection
      for (uint32_t x = 0; x < a.k; ++x) {
        const uint32_t q = s_ord[x];
        if (!fits(q) || ((pos >> q) & 1ull)) continue;
        bq = q;
        break;
      }
      bg = base;
    }
    a.moves[v] = bq;
    a.gains[v] = bq == NONE ? 0 : bg;
  }
  // list slots in this XCD's sub-list: one atomic per workgroup, then wave and lane offsets
  const bool to_list = v < n && !small;
  __shared__ uint32_t s_cnt[4], s_base;
  const uint32_t l = lane_id(), wv = threadIdx.x / 64, xcd = split ? blockIdx.x % XCDS : 0u;
  const uint64_t ml = __ballot(to_list);
  if (l == 0) s_cnt[wv] = (uint32_t)__popcll(ml);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t t = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    s_base = t ? atomicAdd(&count[xcd], t) : 0u;
  }
  __syncthreads();
  uint32_t base = (split ? xcd_list_base(xcd, gridDim.x) : 0u) + s_base;
  for (uint32_t x = 0; x < wv; ++x) base += s_cnt[x];
  if (to_list) list[base + lanes_below(ml)] = v;
}

// Lane q holds partition q's connection; the wave picks the move (its lanes' per-launch
// constants as in the mask kernels below) and lane 0 writes it.
__device__ __forceinline__ void mask_pick(const GainArgs& a, uint32_t v, uint32_t p, uint32_t s, uint32_t l, bool qvalid,
                                          uint32_t psz, uint32_t hq, uint32_t ord, int64_t base, int64_t conn) {
  // best adjacent target: positive connection, room in the target; ties by the hash
  const bool ok = qvalid && l != p && (uint32_t)(psz + s - s / a.discount) <= a.omega;
  int64_t bg = base + conn;
  uint32_t bh = hq, bq = (ok && conn > 0) ? l : NONE;
  wave_best(bg, bh, bq);
  if (bq == NONE) {  // first valid partition in hash order without a positive connection
    const uint32_t qi = ord == NONE ? 0u : ord;
    const uint32_t pq = __shfl(psz, (int)qi, 64);
    const int64_t cq = __shfl(conn, (int)qi, 64);
    const bool fit = ord != NONE && qi != p && (uint32_t)(pq + s - s / a.discount) <= a.omega && cq <= 0;
    const uint64_t fm = __ballot(fit);
    bq = fm ? __shfl(ord, __ffsll((long long)fm) - 1, 64) : NONE;
    bg = base;
  }
  if (l == 0) {
    a.moves[v] = bq;
    a.gains[v] = bq == NONE ? 0 : bg;
  }
}
// Touching hedges jb0, jb0 + stride, ... (64 at a time) of a node in partition p: lane q's
// connection, and the lane's shares of the node's hedge weight and of its saving.  Each chunk's
// masks and weights are staged in the wave's LDS rows (s_lo, s_hi, s_wt: 64 entries each) and
// read back as broadcasts: lane q reads only the mask word that holds bit q.  While the chunk's
// weights are below 2^26 their sum fits 32 bits, and the lane adds in 32 bits.
__device__ __forceinline__ void mask_accumulate(const GainArgs& a, const uint4* __restrict__ hm,
                                                const float* __restrict__ w, uint64_t jb0, uint64_t t1,
                                                uint64_t stride, uint32_t p, uint32_t l, int64_t& conn, int64_t& wl,
                                                int64_t& sl, uint32_t* s_lo, uint32_t* s_hi, uint32_t* s_wt) {
  const uint32_t* col = l < 32 ? s_lo : s_hi;  // the mask word with this lane's partition bit
  const uint32_t sh = l & 31u;
  for (uint64_t jb = jb0; jb < t1; jb += stride) {
    const uint64_t j = jb + l;
    uint32_t mlo = 0, mhi = 0, we = 0;
    if (j < t1) {
      const uint32_t e = a.tch[j];
      const uint4 r = hm[e];
      we = fixed_weight(w[e]);
      mlo = r.x;
      mhi = r.y;
      const uint64_t sg = ((uint64_t)r.w << 32) | r.z;
      wl += we;
      if ((sg >> p) & 1ull) sl += we;
    }
    s_lo[l] = mlo;
    s_hi[l] = mhi;
    s_wt[l] = we;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint32_t cnt = (uint32_t)min<uint64_t>(64, t1 - jb);
    if (__ballot(we >= (1u << 26)) == 0) {  // uniform
      uint32_t c32 = 0;
      for (uint32_t i = 0; i < cnt; ++i) c32 += s_wt[i] & (0u - ((col[i] >> sh) & 1u));
      conn += c32;
    } else {
      for (uint32_t i = 0; i < cnt; ++i) conn += ((col[i] >> sh) & 1u) ? (int64_t)s_wt[i] : 0;
    }
    __builtin_amdgcn_wave_barrier();  // the rows are read before the next chunk is staged
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
}

// One wave per listed node, lane q holding partition q's connection.  The wave loads 64
// touching hedges' records at a time; each record is then broadcast and every lane adds the
// weight when its partition's bit is set.  psizes and the hash order are per-lane constants
// of the launch.  Persistent: the workgroups of XCD x (at least one: the grid is at least
// XCDS) stride over sub-list x, or all of them over the one list.
__global__ void __launch_bounds__(256) k_gains_mask(GainArgs a, const uint4* __restrict__ hm,
                                                    const float* __restrict__ w, const uint32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ count, uint32_t route_grid) {
  __shared__ uint32_t s_rec[4][3][64];  // per wave: a chunk's mask words and weights
  const uint32_t l = lane_id(), wv = threadIdx.x / 64;
  const bool qvalid = l < a.k;
  const uint32_t psz = qvalid ? a.psizes[l] : 0u;
  const uint32_t hq = mix32(l + a.repeat);
  const uint32_t ord = qvalid ? a.order[l] : NONE;  // lane i: the i-th partition in hash order
  const SubList sl_ = sub_list(list, count, route_grid);
  uint64_t bytes = 0;
  for (uint32_t idx = sl_.first * 4 + wv; idx < sl_.count; idx += sl_.step * 4) {
    const uint32_t v = sl_.list[idx];
    const uint32_t p = a.part[v], s = a.sizes[v];
    const uint64_t t0 = a.toff[v], t1 = a.toff[v + 1];
    int64_t conn = 0, wl = 0, sl = 0;
    mask_accumulate(a, hm, w, t0, t1, 64, p, l, conn, wl, sl, s_rec[wv][0], s_rec[wv][1], s_rec[wv][2]);
    mask_pick(a, v, p, s, l, qvalid, psz, hq, ord, wave_sum(sl - wl), conn);  // one reduction: the sums are exact
    bytes += (l == 0 ? 36 : 0) + 24 * (uint64_t)((t1 - t0 + 63 - l) / 64);
  }
  if (a.prof_bytes) {
    bytes = wave_sum(bytes);
    if (l == 0 && bytes) atomicAdd(&a.prof_bytes[blockIdx.x % PROF_SLOTS], (unsigned long long)bytes);
  }
}
// Nodes of thousands of touching hedges (the coarse levels): a wave per node would leave most
// of the chip idle, and the heaviest node would set the launch's time.  Here W waves share a
// node, wave u taking chunks u, u + W, ... of 64 hedges, and each lane adds its hedge's weight
// to the LDS connection of every partition in its mask other than the node's own (most hedges
// of a coarse node lie inside its partition and add nothing), instead of broadcasting every
// record to all lanes.  The chunk after the current one is loaded while the current one is
// added.  The sums are integers, so their order does not matter.
// SMALLK: k <= MASK_SMALL_K (connections summed in registers); the other instance holds no
// such registers, which leaves room for more waves per SIMD.
template <uint32_t W, bool SMALLK>
__global__ void __launch_bounds__(64 * W) k_gains_mask_wide(GainArgs a, const uint4* __restrict__ hm,
                                                            const float* __restrict__ w,
                                                            const uint32_t* __restrict__ list,
                                                            const uint32_t* __restrict__ count, uint32_t route_grid) {
  __shared__ unsigned long long s_conn[MASK_K];
  __shared__ int64_t s_base[W];
  const uint32_t l = lane_id(), wv = threadIdx.x / 64;
  const bool qvalid = l < a.k;
  const uint32_t psz = qvalid ? a.psizes[l] : 0u;
  const uint32_t hq = mix32(l + a.repeat);
  const uint32_t ord = qvalid ? a.order[l] : NONE;
  const SubList sl_ = sub_list(list, count, route_grid);
  if (threadIdx.x < MASK_K) s_conn[threadIdx.x] = 0;
  __syncthreads();
  uint64_t bytes = 0;
  for (uint32_t idx = sl_.first; idx < sl_.count; idx += sl_.step) {  // uniform over the workgroup
    const uint32_t v = sl_.list[idx];
    const uint32_t p = a.part[v], s = a.sizes[v];
    const uint64_t t0 = a.toff[v], t1 = a.toff[v + 1];
    const uint64_t stride = 64 * W;
    const uint64_t notp = ~(1ull << p);
    int64_t wl = 0, sl = 0;
    // with at most MASK_SMALL_K partitions every lane sums its hedges' connections in registers
    // (LDS atomics would all meet on the same few slots) and the wave sums them at the end
    constexpr bool small_k = SMALLK;
    constexpr uint32_t NREG = SMALLK ? MASK_SMALL_K : 1;
    int64_t creg[NREG];
#pragma unroll
    for (uint32_t q = 0; q < NREG; ++q) creg[q] = 0;
    // lane l walks entries j, j + stride, ...: the list entry two chunks ahead and the record
    // one chunk ahead are in flight while the current record is added
    uint64_t j = t0 + 64 * wv + l;
    uint32_t e0 = j < t1 ? a.tch[j] : NONE, e1 = j + stride < t1 ? a.tch[j + stride] : NONE;
    uint4 r0 = make_uint4(0, 0, 0, 0);
    float w0 = 0.0f;
    if (e0 != NONE) {
      r0 = hm[e0];
      w0 = w[e0];
    }
    while (e0 != NONE) {
      const uint32_t e2 = j + 2 * stride < t1 ? a.tch[j + 2 * stride] : NONE;
      uint4 r1 = make_uint4(0, 0, 0, 0);
      float w1 = 0.0f;
      if (e1 != NONE) {
        r1 = hm[e1];
        w1 = w[e1];
      }
      const uint32_t we = fixed_weight(w0);
      const uint64_t sg = ((uint64_t)r0.w << 32) | r0.z;
      wl += we;
      if ((sg >> p) & 1ull) sl += we;
      uint64_t m = (((uint64_t)r0.y << 32) | r0.x) & notp;
      if constexpr (small_k) {
#pragma unroll
        for (uint32_t q = 0; q < NREG; ++q) creg[q] += ((m >> q) & 1ull) ? (int64_t)we : 0;
      } else if (we) {
        while (m) {
          const uint32_t q = (uint32_t)__ffsll((long long)m) - 1;
          m &= m - 1;
          atomicAdd(&s_conn[q], (unsigned long long)we);
        }
      }
      e0 = e1;
      e1 = e2;
      r0 = r1;
      w0 = w1;
      j += stride;
    }
    const int64_t b = wave_sum(sl - wl);
    if (l == 0) s_base[wv] = b;
    if constexpr (small_k)
      for (uint32_t q = 0; q < a.k; ++q) {  // a.k <= NREG
        const int64_t cq = wave_sum(creg[q]);
        if (l == 0 && cq) atomicAdd(&s_conn[q], (unsigned long long)cq);
      }
    __syncthreads();
    if (wv == 0) {
      int64_t bs = 0;
#pragma unroll
      for (uint32_t u = 0; u < W; ++u) bs += s_base[u];
      const int64_t cs = l == p ? 0 : (int64_t)s_conn[l];  // the own partition's sum is not kept
      s_conn[l] = 0;
      mask_pick(a, v, p, s, l, qvalid, psz, hq, ord, bs, cs);
    }
    __syncthreads();  // the sums are cleared before the next node adds to them
    const uint64_t d = t1 - t0, mine = d > 64 * wv + l ? (d - 64 * wv - l + 64 * W - 1) / (64 * W) : 0;
    bytes += (threadIdx.x == 0 ? 36 : 0) + 24 * mine;
  }
  if (a.prof_bytes) {
    bytes = wave_sum(bytes);
    if (l == 0 && bytes) atomicAdd(&a.prof_bytes[blockIdx.x % PROF_SLOTS], (unsigned long long)bytes);
  }
}

// ---- Jacobi label propagation (device initial partitioner) ------------------------------
// movers: positive-gain proposals, keyed for a sort by gain descending (ties: node id)
__global__ void k_jacobi_movers(const uint32_t* __restrict__ moves, const int64_t* __restrict__ gains, uint32_t n,
                                uint32_t* __restrict__ flag) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) flag[v] = (moves[v] != NONE && gains[v] > 0) ? 1u : 0u;
}
__global__ void k_jacobi_keys(const uint32_t* __restrict__ flag, const uint64_t* __restrict__ pos,
                              const int64_t* __restrict__ gains, uint32_t n, uint64_t* __restrict__ keys,
                              uint32_t* __restrict__ vals) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n || !flag[v]) return;
  keys[pos[v]] = ~ordered_i64(gains[v]);
  vals[pos[v]] = v;
}
__global__ void k_jacobi_dst(const uint32_t* __restrict__ vals, const uint32_t* __restrict__ moves,
                             const uint32_t* __restrict__ sizes, uint32_t m, uint32_t* __restrict__ dst,
                             int64_t* __restrict__ sz) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  dst[i] = moves[vals[i]];
  sz[i] = sizes[vals[i]];
}
__global__ void k_jacobi_sizes(const uint32_t* __restrict__ vals, const uint32_t* __restrict__ sizes, uint32_t m,
                               int64_t* __restrict__ sz) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) sz[i] = sizes[vals[i]];
}
// Sorted by (target asc, gain desc, node asc): a move applies when everything entering its
// target up to it still fits next to the target's size before the round (moves leaving the
// target are not counted on).
__global__ void __launch_bounds__(256) k_jacobi_apply(const uint32_t* __restrict__ dst,
                                                      const uint32_t* __restrict__ vals,
                                                      const int64_t* __restrict__ incl,
                                                      const uint64_t* __restrict__ doff, uint32_t m,
                                                      const uint32_t* __restrict__ sizes,
                                                      const uint32_t* __restrict__ before, uint32_t omega,
                                                      uint32_t* __restrict__ part, uint32_t* __restrict__ psizes,
                                                      uint32_t* __restrict__ applied) {
  __shared__ uint32_t s_part[4];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t ok = 0;
  if (i < m) {
    const uint32_t q = dst[i];
    const uint64_t s0 = doff[q];
    const int64_t entering = incl[i] - (s0 ? incl[s0 - 1] : 0);
    if ((int64_t)before[q] + entering <= (int64_t)omega) {
      const uint32_t v = vals[i];
      const uint32_t sv = sizes[v];
      atomicSub(&psizes[part[v]], sv);
      atomicAdd(&psizes[q], sv);
      part[v] = q;
      ok = 1;
    }
  }
  const uint32_t c = wave_sum(ok);
  if (lane_id() == 0) s_part[threadIdx.x / 64] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t t = s_part[0] + s_part[1] + s_part[2] + s_part[3];
    if (t) atomicAdd(applied, t);
  }
}

// partitions in descending mix32(q + repeat) order (mix32 is a bijection: no ties)
__global__ void k_order_keys(uint32_t k, uint32_t repeat, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= k) return;
  keys[q] = ~mix32(q + repeat);
  vals[q] = q;
}

__global__ void k_compact_u64(const uint32_t* __restrict__ flag, const uint64_t* __restrict__ pos, uint32_t n,
                              const uint64_t* __restrict__ in, uint64_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) out[pos[i]] = in[i];
}
// the small-hedge CSR in first-pin order: sorted keys (first pin << 32 | hedge) -> lengths,
// then pins and weights at the scanned offsets
// The small-hedge CSR: hedge i (hedge keys[i], or i without keys) keeps its pins from a
// multiple of 4, padded to a multiple of 4 with its last pin, so a thread loads them as 16-byte
// vectors; its record is start << 5 | length (lengths 2..16).
__device__ __forceinline__ uint32_t small_hedge_id(const uint64_t* keys, uint32_t i) {
  return keys ? (uint32_t)keys[i] : i;
}
__global__ void k_small_lens(const uint64_t* __restrict__ keys, uint32_t m, const uint64_t* __restrict__ hoff,
                             uint64_t* __restrict__ len) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t e = small_hedge_id(keys, i);
  len[i] = (hoff[e + 1] - hoff[e] + 3) & ~3ull;
}
__global__ void k_small_fill(const uint64_t* __restrict__ keys, uint32_t m, const uint64_t* __restrict__ hoff,
                             const uint32_t* __restrict__ hedges, const float* __restrict__ w,
                             uint64_t* __restrict__ rec, uint32_t* __restrict__ pins, float* __restrict__ sw) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t e = small_hedge_id(keys, i);
  const uint64_t s = hoff[e], len = hoff[e + 1] - s, o = rec[i];  // rec holds the start until now
  for (uint64_t j = 0; j < len; ++j) pins[o + j] = hedges[s + j];
  for (uint64_t j = len; j < ((len + 3) & ~3ull); ++j) pins[o + j] = hedges[s + len - 1];
  sw[i] = w[e];
  rec[i] = (o << 5) | len;
}
__global__ void k_compact(const uint32_t* __restrict__ flag, const uint64_t* __restrict__ pos, uint32_t n,
                          uint32_t* __restrict__ out) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n && flag[v]) out[pos[v]] = v;
}


__global__ void k_split(const uint32_t* __restrict__ flag, const uint64_t* __restrict__ pos, uint32_t n,
                        uint32_t* __restrict__ a, uint32_t* __restrict__ b) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  if (flag[v]) a[pos[v]] = v;
  else b[v - pos[v]] = v;
}

__global__ void k_heavy_tabcap(const uint32_t* __restrict__ list, const uint64_t* __restrict__ bound, uint32_t m,
                               uint32_t k, uint64_t* __restrict__ cap) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  uint64_t b = bound[list[i]];
  if (b > k) b = k;
  b = 2 * b + 64;
  uint64_t c = 1;
  while (c < b) c <<= 1;
  cap[i] = c;
}

// ---- 3. ranking -------------------------------------------------------------------------
__global__ void k_rank_keys_gain(const uint32_t* __restrict__ moves, const int64_t* __restrict__ gains, uint32_t n,
                                 uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  keys[v] = moves[v] == NONE ? 0ull : ~ordered_i64(gains[v]);
  vals[v] = v;
}
__global__ void k_invalid_keys(const uint32_t* __restrict__ moves, const uint32_t* __restrict__ vals, uint32_t n,
                               uint32_t* __restrict__ keys) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keys[i] = moves[vals[i]] == NONE ? 1u : 0u;
}
__global__ void k_scatter_ranks(const uint32_t* __restrict__ vals, uint32_t n, uint32_t* __restrict__ ranks) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) ranks[vals[r]] = r;
}
// grid-stride count, reduced per workgroup: one atomic per workgroup
__global__ void __launch_bounds__(256) k_count_valid(const uint32_t* __restrict__ moves, uint32_t n, uint32_t* cnt) {
  __shared__ uint32_t s_part[4];
  uint32_t c = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    c += moves[v] != NONE ? 1u : 0u;
  c = wave_sum(c);
  if (lane_id() == 0) s_part[threadIdx.x / 64] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t t = s_part[0] + s_part[1] + s_part[2] + s_part[3];
    if (t) atomicAdd(cnt, t);
  }
}

// zeroes three arrays of n entries in one launch (instead of three fills)
__global__ void k_zero3(unsigned long long* __restrict__ a, uint32_t* __restrict__ b, uint32_t* __restrict__ c,
                        uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  a[i] = 0ull;
  b[i] = 0u;
  c[i] = 0u;
}

// chaining: edges in sorted space j (by source partition asc, valid first, gain desc)
struct ChainArrays {
  const uint32_t* orig;  // j -> node
  uint32_t* src;
  uint32_t* dst;
  uint32_t* sz;
  int64_t* w;
  uint32_t* ok;
  uint32_t* ob;
  uint32_t* oe;
  uint32_t* next;
  uint32_t* prev;
  unsigned long long* claim;
  uint32_t* progress;  // progress[it] != 0: iteration it committed a link
  uint4* pk;           // per entry: gain (lo, hi), size, valid -- what a proposal scan reads
  uint32_t* from;      // per entry: where its next proposal scan starts
};
// An iteration after one that committed nothing would repeat it exactly: skip it.
__device__ __forceinline__ bool chain_stalled(const ChainArrays& c, int it) {
  return it > 0 && c.progress[it - 1] == 0;
}

__global__ void k_chain_sortkeys(const uint32_t* __restrict__ part, const uint32_t* __restrict__ moves,
                                 const int64_t* __restrict__ gains, uint32_t n, uint64_t* __restrict__ k1,
                                 uint32_t* __restrict__ vals) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  k1[v] = moves[v] == NONE ? 0ull : ~ordered_i64(gains[v]);
  vals[v] = v;
}
// second (more significant) pass key: (source partition << 1) | invalid
__global__ void k_chain_sortkeys2(const uint32_t* __restrict__ part, const uint32_t* __restrict__ moves,
                                  const uint32_t* __restrict__ vals, uint32_t n, uint32_t* __restrict__ k2) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t v = vals[i];
  k2[i] = (part[v] << 1) | (moves[v] == NONE ? 1u : 0u);
}
__global__ void k_chain_gather(ChainArrays c, const uint32_t* __restrict__ part, const uint32_t* __restrict__ moves,
                               const int64_t* __restrict__ gains, const uint32_t* __restrict__ sizes, uint32_t n) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t v = c.orig[j];
  c.src[j] = part[v];
  c.dst[j] = moves[v];
  c.sz[j] = sizes[v];
  c.w[j] = gains[v];
  c.ok[j] = moves[v] != NONE;
  c.pk[j] = make_uint4((uint32_t)(uint64_t)gains[v], (uint32_t)((uint64_t)gains[v] >> 32), sizes[v],
                       moves[v] != NONE ? 1u : 0u);
  c.next[j] = NONE;
  c.prev[j] = NONE;
  c.claim[j] = 0ull;
}
__global__ void k_chain_ranges(ChainArrays c, const uint64_t* __restrict__ poff, uint32_t n) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (!c.ok[j]) {
    c.ob[j] = 0;
    c.oe[j] = 0;
    return;
  }
  const uint32_t q = c.dst[j];
  c.ob[j] = (uint32_t)poff[q];
  c.oe[j] = (uint32_t)poff[q + 1];
  c.from[j] = c.ob[j];
}
__device__ __forceinline__ int size_bucket(uint32_t s) { return s == 0 ? 0 : 31 - __clz(s); }

// Each open move proposes to the best free move out of its destination partition among the
// first 256 entries of that partition's list.  The list is sorted by gain (descending, valid
// moves first), so once a choice exists the scan stops at the first lower gain or invalid
// entry: nothing later can beat it.  An entry that cannot be chosen (invalid, the proposer
// itself, already linked, or of a size bucket too far) never can again, so the next scan
// starts at the first entry this one could choose.
// Claims, in the same launch: the proposer with the smallest size difference, then the lowest
// index, wins.  Proposals pile onto few targets (the heads of the partition lists) but the
// lanes of a wave rarely share one, so the workgroup first merges its claims per target in an
// LDS table and then issues one global atomic per distinct target.  Claims start at zero: the
// buffer is cleared when it is filled, and the commit clears every claim it reads.
constexpr uint32_t RESOLVE_BLOCK = 1024;
constexpr uint32_t RESOLVE_SLOTS = 2048;  // > 2 x the workgroup's claims: the table never fills
constexpr uint32_t PROPOSE_BATCH = 6;     // list entries per load round (60 VGPRs: two workgroups per CU)
__global__ void __launch_bounds__(RESOLVE_BLOCK) k_chain_propose(ChainArrays c, uint32_t n, int it) {
  if (chain_stalled(c, it)) return;  // uniform
  __shared__ uint32_t s_tgt[RESOLVE_SLOTS];
  __shared__ unsigned long long s_key[RESOLVE_SLOTS];
  for (uint32_t t = threadIdx.x; t < RESOLVE_SLOTS; t += RESOLVE_BLOCK) {
    s_tgt[t] = NONE;
    s_key[t] = 0ull;
  }
  const uint32_t i = blockIdx.x * RESOLVE_BLOCK + threadIdx.x;
  uint32_t choice = NONE;
  if (i < n && c.next[i] == NONE && c.ob[i] < c.oe[i]) {
    const uint32_t si = c.sz[i];
    const int bi = size_bucket(si);
    const uint32_t lim = min(c.oe[i], c.ob[i] + 256u);
    const uint32_t j0 = max(c.from[i], c.ob[i]);
    uint32_t first = lim;  // first entry that could be chosen (scan end when none)
    int64_t bw = 0;
    uint32_t bd = 0;
    bool done = false;
    // PROPOSE_BATCH entries are loaded before any is examined, so a scan waits on one memory
    // round trip per batch instead of one per entry
    for (uint32_t jb = j0; jb < lim && !done; jb += PROPOSE_BATCH) {
      uint4 e[PROPOSE_BATCH];
      uint32_t pv[PROPOSE_BATCH];
#pragma unroll
      for (uint32_t u = 0; u < PROPOSE_BATCH; ++u) {  // unconditional loads (clamped index): no branch per load
        const uint32_t jj = min(jb + u, lim - 1);
        e[u] = c.pk[jj];
        pv[u] = c.prev[jj];
        if (jb + u >= lim) e[u].w = 0;
      }
#pragma unroll
      for (uint32_t u = 0; u < PROPOSE_BATCH; ++u) {
        const uint32_t j = jb + u;
        if (!e[u].w) {  // invalid entry, or past the scan limit
          if (choice == NONE) first = min(j, lim);
          done = true;
          break;
        }
        const int64_t wj = (int64_t)(((uint64_t)e[u].y << 32) | e[u].x);
        if (choice != NONE && wj < bw) {
          done = true;
          break;
        }
        if (j == i || pv[u] != NONE) continue;
        const uint32_t sj = e[u].z;
        if (abs(bi - size_bucket(sj)) > 1) continue;
        const uint32_t d = si > sj ? si - sj : sj - si;
        if (choice == NONE) first = j;
        if (choice == NONE || wj > bw || (wj == bw && d < bd)) {
          choice = j;
          bw = wj;
          bd = d;
          if (d == 0) {  // later entries gain no more and cannot have a smaller difference
            done = true;
            break;
          }
        }
      }
    }
    c.from[i] = first;
  }
  __syncthreads();  // the table is cleared
  if (choice != NONE && c.prev[choice] == NONE) {
    const uint32_t d = c.sz[i] > c.sz[choice] ? c.sz[i] - c.sz[choice] : c.sz[choice] - c.sz[i];
    const unsigned long long key = ((unsigned long long)(0xFFFFFFFFu - d) << 32) | (0xFFFFFFFFu - i);
    uint32_t h = slot_hash(choice) & (RESOLVE_SLOTS - 1);
    for (;;) {
      const uint32_t old = atomicCAS(&s_tgt[h], NONE, choice);
      if (old == NONE || old == choice) break;
      h = (h + 1) & (RESOLVE_SLOTS - 1);
    }
    atomicMax(&s_key[h], key);
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < RESOLVE_SLOTS; t += RESOLVE_BLOCK)
    if (s_tgt[t] != NONE) atomicMax(&c.claim[s_tgt[t]], s_key[t]);
}
__global__ void k_chain_commit(ChainArrays c, uint32_t n, int it) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || chain_stalled(c, it)) return;
  const unsigned long long k = c.claim[s];
  if (k == 0ull) return;
  c.claim[s] = 0ull;  // the next iteration's claims start at zero
  if (c.prev[s] != NONE) return;
  const uint32_t p = 0xFFFFFFFFu - (uint32_t)(k & 0xFFFFFFFFull);
  c.prev[s] = p;
  if (c.next[p] == NONE) c.next[p] = s;
  c.progress[it] = 1u;
}
// comp / pos: each move's component (its chain's head) and its distance from the head; a
// chain whose head is not found within 256 steps (a cycle, or a longer chain) is keyed by
// its lowest index at position 0 and raises *odd.  csize counts each component's moves.
__global__ void k_chain_comp(ChainArrays c, uint32_t n, uint32_t* __restrict__ comp, uint32_t* __restrict__ pos,
                             uint32_t* __restrict__ cflag, long long* __restrict__ cw,
                             uint32_t* __restrict__ cinv, uint32_t* __restrict__ csize, uint32_t* __restrict__ odd) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t cur = i, mn = i, cc = NONE, ps = 0;
  for (uint32_t d = 0; d < 256; ++d) {
    const uint32_t p = c.prev[cur];
    if (p == NONE) {
      cc = cur;
      ps = d;
      break;
    }
    cur = p;
    mn = cur < mn ? cur : mn;
  }
  if (cc == NONE) {
    cc = mn;
    ps = 0;
    *odd = 1u;
  }
  comp[i] = cc;
  pos[i] = ps;
  cflag[cc] = 1u;
  atomicAdd(&csize[cc], 1u);
  if (c.ok[i]) atomicAdd((unsigned long long*)&cw[cc], (unsigned long long)c.w[i]);  // two's complement sum
  else atomicOr(&cinv[cc], 1u);
}
__global__ void k_comp_list(const uint32_t* __restrict__ cflag, const uint64_t* __restrict__ cpos, uint32_t n,
                            const long long* __restrict__ cw, const uint32_t* __restrict__ cinv,
                            uint32_t* __restrict__ clist, uint64_t* __restrict__ ckey, uint32_t* __restrict__ cinvk) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || !cflag[j]) return;
  const uint64_t o = cpos[j];
  clist[o] = j;
  ckey[o] = ~ordered_i64((int64_t)cw[j]);
  cinvk[o] = cinv[j];
}
__global__ void k_comp_rank(const uint32_t* __restrict__ clist, uint32_t m, uint32_t* __restrict__ crank) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < m) crank[clist[r]] = r;
}
// sizes of the components in rank order
__global__ void k_comp_sizes(const uint32_t* __restrict__ clist, const uint32_t* __restrict__ csize, uint32_t m,
                             uint32_t* __restrict__ out) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < m) out[r] = csize[clist[r]];
}
// Every chain has its head: a move's rank is its component's start (the sizes of the
// components ranked before it) plus its distance from the head -- the position the stable
// sort by (component rank, distance) would give it, distances being unique within a chain.
__global__ void k_chain_direct_ranks(const uint32_t* __restrict__ comp, const uint32_t* __restrict__ pos,
                                     const uint32_t* __restrict__ crank, const uint64_t* __restrict__ start,
                                     const uint32_t* __restrict__ orig, uint32_t n, uint32_t* __restrict__ ranks) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) ranks[orig[j]] = (uint32_t)(start[crank[comp[j]]] + pos[j]);
}
__global__ void k_pack2(const uint64_t* __restrict__ a, const uint32_t* __restrict__ b, uint64_t* __restrict__ out) {
  out[0] = a[0];
  out[1] = b[0];
}
__global__ void k_edge_keys(const uint32_t* __restrict__ comp, const uint32_t* __restrict__ pos,
                            const uint32_t* __restrict__ crank, uint32_t n, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ vals) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  keys[j] = ((uint64_t)crank[comp[j]] << 8) | pos[j];
  vals[j] = j;
}
__global__ void k_edge_ranks(const uint32_t* __restrict__ sorted_j, const uint32_t* __restrict__ orig, uint32_t n,
                             uint32_t* __restrict__ ranks) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) ranks[orig[sorted_j[r]]] = r;
}

// ---- 4. cascade (hedge-centric) ------------------------------------------------------------
struct CascadeArgs {
  const uint32_t* hedges;
  const uint64_t* hoff;
  const float* w;
  // per node (partition, move target or NONE, rank): one 8-byte gather per pin when every
  // partition id fits 16 bits (pm2: partition | target << 16, target NONE as 0xFFFF; rank),
  // otherwise one 16-byte gather (pm: partition, target, rank, 0)
  const uint4* pm;
  const uint2* pm2;
  int incc;
  unsigned long long* seq;
};
constexpr uint32_t PM2_MAX_K = 0xFFFFu;  // partition ids < 0xFFFF leave 0xFFFF for "no move"

// (partition, move target or NONE, rank) of node x
__device__ __forceinline__ uint3 pm_at(const CascadeArgs& a, uint32_t x) {
  if (a.pm2) {
    const uint2 r = a.pm2[x];
    const uint32_t q = r.x >> 16;
    return make_uint3(r.x & 0xFFFFu, q == 0xFFFFu ? NONE : q, r.y);
  }
  const uint4 r = a.pm[x];
  return make_uint3(r.x, r.y, r.z);
}

__global__ void k_pack_nodes(const uint32_t* __restrict__ part, const uint32_t* __restrict__ moves,
                             const uint32_t* __restrict__ ranks, uint32_t n, uint4* __restrict__ pm) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) pm[v] = make_uint4(part[v], moves[v], ranks[v], 0u);
}
__global__ void k_pack_nodes2(const uint32_t* __restrict__ part, const uint32_t* __restrict__ moves,
                              const uint32_t* __restrict__ ranks, uint32_t n, uint2* __restrict__ pm2) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint32_t q = moves[v] == NONE ? 0xFFFFu : moves[v];
  pm2[v] = make_uint2(part[v] | (q << 16), ranks[v]);
}

__device__ __forceinline__ int64_t cascade_term(int incc, int64_t W, int64_t cp, int64_t cq) {
  int64_t t = 0;
  if (incc) {
    if (cp >= 1) t += (int64_t)((uint64_t)W / ((uint64_t)cp * (uint64_t)cp));
  } else if (cp == 1) {
    t += W;
  }
  if (cq == 0) t -= W;
  return t;
}

// Cascade classes of a level's hedges: 2..CAS_SMALL pins (the thread-per-hedge kernel; the
// key orders them by their first pin), the sorted kernels (up to CAS_SORT_S or sort_max
// pins) and the direct LDS/HBM kernel.  A hedge of one pin adds nothing: its pin finds itself
// alone in p and nobody in q, so the two terms of cascade_term cancel.
__global__ void k_cascade_classes(const uint32_t* __restrict__ hedges, const uint64_t* __restrict__ hoff, uint32_t E,
                                  uint32_t sort_max, uint32_t* __restrict__ f_small, uint64_t* __restrict__ k_small,
                                  uint32_t* __restrict__ f_s, uint32_t* __restrict__ f_m, uint32_t* __restrict__ f_big) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const uint64_t s = hoff[e], hl = hoff[e + 1] - s;
  const bool small = hl >= 2 && hl <= CAS_SMALL;
  const bool sorted = hl > CAS_SMALL && hl <= sort_max;
  f_small[e] = small ? 1u : 0u;
  k_small[e] = small ? ((uint64_t)hedges[s] << 32) | e : 0ull;
  f_s[e] = (sorted && hl <= CAS_SORT_S) ? 1u : 0u;
  f_m[e] = (sorted && hl > CAS_SORT_S) ? 1u : 0u;
  f_big[e] = (hl > CAS_SMALL && !sorted) ? 1u : 0u;
}

// The level's hedges of 2..CAS_SMALL pins as their own CSR, in order of their first pin (built
// once per level): offsets, pins and weights are read coalesced, and nearby threads gather the
// records of nearby nodes.
struct SmallHedges {
  const uint64_t* rec;  // start << 5 | length; starts are multiples of 4
  const uint32_t* pins;
  const float* w;
  uint32_t m;
};

// Hedges of 2..CAS_SMALL pins, one thread each, each XCD given a contiguous range of the
// first-pin order so that the records its threads gather meet in one L2.
// P16 (packed records, k < PM2_MAX_K): a pin's partition and target stay packed in one
// register (partition | target << 16, 0xFFFF = none), so the pins take 32 registers instead
// of 48 and more waves fit on a SIMD.
template <bool P16>
__global__ void k_cascade_small(CascadeArgs a, SmallHedges h) {
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= h.m) return;
  const uint64_t rec = h.rec[i];
  const uint64_t s = rec >> 5;
  const uint32_t len = (uint32_t)(rec & 31u);
  const uint4* p4 = reinterpret_cast<const uint4*>(h.pins + s);  // 16-byte aligned: s % 4 == 0
  // per pin: partition (with P16 also the target), target, rank
  uint32_t pt[CAS_SMALL], mv[P16 ? 1 : CAS_SMALL], rk[CAS_SMALL];
  // a pin's partition, its target (NONE when it does not move) and whether it moves; entries
  // past the hedge's end are non-moving pins of no partition (0xFFFF or NONE: no real id)
  auto part_of = [&](uint32_t u) {
    if constexpr (P16) return pt[u] & 0xFFFFu;
    else return pt[u];
  };
  auto move_of = [&](uint32_t u) {
    if constexpr (P16) return pt[u] >> 16;
    else return mv[u];
  };
  auto moves = [&](uint32_t u) {
    if constexpr (P16) return (pt[u] >> 16) != 0xFFFFu;
    else return mv[u] != NONE;
  };
#pragma unroll
  for (uint32_t u = 0; u < CAS_SMALL; ++u) {  // entries no batch fills (u >= 4 when len <= 4)
    pt[u] = P16 ? 0xFFFFFFFFu : NONE;
    if constexpr (!P16) mv[u] = NONE;
    rk[u] = 0;
  }
  // The pins come as 16-byte vectors (the padding repeats the last pin) and their records in
  // batches of unconditional loads, so a hedge costs two memory round trips per batch instead
  // of two per pin; entries past the end are marked as non-moving pins of no partition.
  auto put = [&](uint32_t u, uint32_t x) {
    const bool in = u < len;
    if constexpr (P16) {
      const uint2 r = a.pm2[x];
      pt[u] = in ? r.x : 0xFFFFFFFFu;
      rk[u] = r.y;
    } else {
      const uint3 q4 = pm_at(a, x);
      pt[u] = in ? q4.x : NONE;
      mv[u] = in ? q4.y : NONE;
      rk[u] = q4.z;
    }
  };
  static_assert(CAS_SMALL == 16, "four vectors of four pins");
  const uint4 v0 = p4[0];
  {
    const uint32_t x[4] = {v0.x, v0.y, v0.z, v0.w};
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) put(u, x[u]);
  }
  if (len > 4) {
    // vectors past the hedge's padded pins are not loaded: their entries repeat pin 0
    const uint4 v1 = p4[1];
    const uint4 v2 = len > 8 ? p4[2] : make_uint4(v0.x, v0.x, v0.x, v0.x);
    const uint4 v3 = len > 12 ? p4[3] : make_uint4(v0.x, v0.x, v0.x, v0.x);
    const uint32_t x[CAS_SMALL - 4] = {v1.x, v1.y, v1.z, v1.w, v2.x, v2.y, v2.z, v2.w, v3.x, v3.y, v3.z, v3.w};
#pragma unroll
    for (uint32_t u = 4; u < CAS_SMALL; ++u) put(u, x[u - 4]);
  }
  bool any = false;
#pragma unroll
  for (uint32_t i = 0; i < CAS_SMALL; ++i) any |= moves(i);
  if (!any) return;
  const int64_t W = (int64_t)fixed_weight(h.w[i]);
  // Static indexes only (the arrays stay in registers); entries past a hedge's end take no
  // part.  The loops stop at the wave's longest hedge, rounded up to a multiple of 4.
  const uint32_t jmax = __ballot(len > 12) ? 16u : __ballot(len > 8) ? 12u : __ballot(len > 4) ? 8u : 4u;
#pragma unroll
  for (uint32_t i = 0; i < CAS_SMALL; ++i) {
    if (i < jmax && moves(i)) {
      const uint32_t p = part_of(i), q = move_of(i), r = rk[i];
      int32_t cp = 0, cq = 0;
#pragma unroll
      for (uint32_t j = 0; j < CAS_SMALL; ++j) {
        if (j < jmax) {
          const bool before = moves(j) && rk[j] < r;         // moved before pin i
          const uint32_t now = before ? move_of(j) : part_of(j);  // its partition when pin i moves
          cp += now == p;
          cq += now == q;
        }
      }
      const int64_t term = cascade_term(a.incc, W, cp, cq);
      if (term) atomicAdd(&a.seq[r], (unsigned long long)term);
    }
  }
}

// Algorithmic bytes of one k_cascade_small launch, counted when that kernel is profiled (a
// separate launch, outside the timed one).  Per hedge of 2..CAS_SMALL pins: its two offsets
// (8 B, shared with its neighbours); every pin id (4 B) and its node's record (8 B packed, 16 B
// otherwise); the weight (4 B) when a pin moves, and one 8-byte cascade term per moving pin.
__global__ __launch_bounds__(256) void k_cascade_small_bytes(CascadeArgs a, SmallHedges h,
                                                             unsigned long long* __restrict__ prof) {
  const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t b = 0;
  if (i0 < h.m) {
    const uint64_t s = h.rec[i0] >> 5, len = h.rec[i0] & 31u, rec = a.pm2 ? 8 : 16;
    uint32_t moving = 0;
    for (uint64_t i = s; i < s + len; ++i) moving += pm_at(a, h.pins[i]).y != NONE;
    b = 8 + (4 + rec) * len + (moving ? 4 + 8 * (uint64_t)moving : 0);
  }
  __shared__ uint64_t part[4];
  const uint64_t ws = wave_sum(b);
  if (lane_id() == 0) part[threadIdx.x / 64] = ws;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicAdd(&prof[blockIdx.x % PROF_SLOTS], (unsigned long long)(part[0] + part[1] + part[2] + part[3]));
}

__global__ void __launch_bounds__(64) k_cascade_large(CascadeArgs a, const uint32_t* __restrict__ list) {
  __shared__ uint32_t s_pt[CAS_LDS], s_mv[CAS_LDS], s_rk[CAS_LDS];
  const uint32_t e = list[blockIdx.x];
  const uint32_t l = lane_id();
  const uint64_t s = a.hoff[e];
  const uint32_t len = (uint32_t)(a.hoff[e + 1] - s);
  const bool in_lds = len <= CAS_LDS;
  bool any = false;
  if (in_lds) {
    for (uint32_t i = l; i < len; i += 64) {
      const uint32_t x = a.hedges[s + i];
      const uint3 q4 = pm_at(a, x);
      s_pt[i] = q4.x;
      s_mv[i] = q4.y;
      s_rk[i] = q4.z;
      any |= s_mv[i] != NONE;
    }
  } else {
    for (uint32_t i = l; i < len; i += 64) any |= pm_at(a, a.hedges[s + i]).y != NONE;
  }
  __syncthreads();
  if (!__ballot(any)) return;
  const int64_t W = (int64_t)fixed_weight(a.w[e]);
  auto PT = [&](uint32_t j) { return in_lds ? s_pt[j] : pm_at(a, a.hedges[s + j]).x; };
  auto MV = [&](uint32_t j) { return in_lds ? s_mv[j] : pm_at(a, a.hedges[s + j]).y; };
  auto RK = [&](uint32_t j) { return in_lds ? s_rk[j] : pm_at(a, a.hedges[s + j]).z; };
  for (uint32_t i = l; i < len; i += 64) {
    const uint32_t q = MV(i);
    if (q == NONE) continue;
    const uint32_t p = PT(i), r = RK(i);
    int64_t cp = 0, cq = 0;
    for (uint32_t j = 0; j < len; ++j) {
      const uint32_t pj = PT(j), mj = MV(j);
      cp += pj == p;
      cq += pj == q;
      if (mj != NONE && RK(j) < r) {
        if (mj == p) cp++;
        else if (mj == q) cq++;
        if (pj == p) cp--;
        else if (pj == q) cq--;
      }
    }
    atomicAdd(&a.seq[r], (unsigned long long)cascade_term(a.incc, W, cp, cq));
  }
}

// Cascade of one hedge through sorted events (hedges of CAS_SMALL < len <= CAS_SORT pins,
// one workgroup per hedge).  The pins' partitions are counted in an LDS hash table; every
// moving pin contributes an "out" event (source, -1) and an "in" event (target, +1) at its
// rank.  Sorted by (partition, rank), a move event's exclusive prefix within its partition,
// plus that partition's pin count, is the pin count of its source just before the move (the
// pin itself included) for the out event and that of its target for the in event: exactly
// c_p and c_q of the cascade term, in O(m log^2 m) for m moving pins instead of
// O(len * m).
// key = partition << (rb + 11) | (rank + 1) << 11 | dir << 10 | local index (dir 1 = in).
__device__ __forceinline__ int32_t wave_seg_scan(int32_t v, bool& flag) {
  const uint32_t l = lane_id();
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const int32_t pv = __shfl_up(v, d, 64);
    const int pf = __shfl_up((int)flag, d, 64);
    if (l >= d) {
      if (!flag) v += pv;
      flag = flag || pf;
    }
  }
  return v;
}

template <uint32_t PINS, uint32_t EVS, uint32_t NT>
__global__ void __launch_bounds__(NT) k_cascade_sorted(CascadeArgs a, const uint32_t* __restrict__ list, int rb) {
  static_assert(EVS >= 2 * PINS && (EVS & (EVS - 1)) == 0, "two event slots per pin, a power of two");
  constexpr uint32_t HS = EVS;  // partition table: at least twice the pins, so it never fills
  __shared__ uint64_t s_ev[EVS];
  __shared__ uint32_t s_hk[HS], s_hc[HS];
  __shared__ uint32_t s_rk[PINS];
  __shared__ int32_t s_cp[PINS], s_cq[PINS];
  __shared__ uint32_t s_moving;
  const uint32_t e = list[blockIdx.x];
  const uint32_t l = lane_id(), tid = threadIdx.x;
  const uint64_t s = a.hoff[e];
  const uint32_t len = (uint32_t)(a.hoff[e + 1] - s);
  if (tid == 0) s_moving = 0;
  for (uint32_t h = tid; h < HS; h += NT) {
    s_hk[h] = NONE;
    s_hc[h] = 0;
  }
  __syncthreads();
  // pin counts per partition, and the two events of each moving pin
  for (uint32_t i = tid; i < len; i += NT) {
    const uint32_t x = a.hedges[s + i];
    const uint3 q4 = pm_at(a, x);
    const uint32_t p = q4.x, q = q4.y;
    uint32_t h = slot_hash(p) & (HS - 1);
    for (;;) {
      const uint32_t old = atomicCAS(&s_hk[h], NONE, p);
      if (old == NONE || old == p) break;
      h = (h + 1) & (HS - 1);
    }
    atomicAdd(&s_hc[h], 1u);
    s_rk[i] = q == NONE ? NONE : q4.z;
    // event slots: one LDS atomic per wave (the lowest moving lane's), not one per moving pin
    const uint64_t mv = __ballot(q != NONE);
    uint32_t o0 = 0;
    const uint32_t lead = mv ? (uint32_t)__ffsll((long long)mv) - 1u : 0u;
    if (mv && l == lead) o0 = atomicAdd(&s_moving, (uint32_t)__popcll(mv));
    o0 = (uint32_t)__builtin_amdgcn_readlane((int)o0, (int)lead);
    if (q != NONE) {
      const uint32_t o = 2 * (o0 + (uint32_t)__popcll(mv & ((1ull << l) - 1ull)));
      const uint64_t r1 = (uint64_t)q4.z + 1;
      s_ev[o] = ((uint64_t)p << (rb + 11)) | (r1 << 11) | i;
      s_ev[o + 1] = ((uint64_t)q << (rb + 11)) | (r1 << 11) | (1u << 10) | i;
    }
  }
  __syncthreads();
  const uint32_t moving = s_moving;
  if (moving == 0) return;
  const uint32_t nev = 2 * moving;
  uint32_t ev_n = 64;
  while (ev_n < nev) ev_n <<= 1;
  for (uint32_t i = nev + tid; i < ev_n; i += NT) s_ev[i] = ~0ull;
  __syncthreads();
  // bitonic sort, ascending
  for (uint32_t k = 2; k <= ev_n; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t t = tid; t < ev_n / 2; t += NT) {
        const uint32_t i = (t / j) * 2 * j + (t % j);
        const uint32_t ixj = i + j;
        const uint64_t x = s_ev[i], y = s_ev[ixj];
        const bool up = (i & k) == 0;
        if ((x > y) == up) {
          s_ev[i] = y;
          s_ev[ixj] = x;
        }
      }
      __syncthreads();
    }
  }
  // segmented (by partition) exclusive prefix of the move deltas, by the first wave; each
  // event adds its partition's pin count (0 for a target partition with no pin here)
  if (tid < 64) {
    const int shift = rb + 11;
    int32_t carry = 0;
    for (uint32_t b = 0; b < nev; b += 64) {
      const uint32_t i = b + l;
      const bool act = i < nev;
      const uint64_t key = act ? s_ev[i] : ~0ull;
      const bool dir_in = (key >> 10) & 1u;
      const int32_t d = !act ? 0 : (dir_in ? 1 : -1);
      bool flag = act && (i == 0 || (s_ev[i - 1] >> shift) != (key >> shift));
      int32_t v = wave_seg_scan(d, flag);
      if (!flag) v += carry;  // segment continues from the previous chunk
      if (act) {
        const uint32_t x = (uint32_t)(key >> shift);
        uint32_t h = slot_hash(x) & (HS - 1), cnt = 0;
        for (;;) {
          const uint32_t kx = s_hk[h];
          if (kx == x) {
            cnt = s_hc[h];
            break;
          }
          if (kx == NONE) break;
          h = (h + 1) & (HS - 1);
        }
        const uint32_t idx = (uint32_t)(key & 1023u);
        if (dir_in) s_cq[idx] = (int32_t)cnt + v - d;
        else s_cp[idx] = (int32_t)cnt + v - d;
      }
      carry = __shfl(v, 63, 64);
    }
  }
  __syncthreads();
  const int64_t W = (int64_t)fixed_weight(a.w[e]);
  for (uint32_t i = tid; i < len; i += NT) {
    const uint32_t r = s_rk[i];
    if (r == NONE) continue;
    atomicAdd(&a.seq[r], (unsigned long long)cascade_term(a.incc, W, s_cp[i], s_cq[i]));
  }
}

// ---- 5. events ----------------------------------------------------------------------------
__global__ void k_size_events(const uint32_t* __restrict__ part, const uint32_t* __restrict__ moves,
                              const uint32_t* __restrict__ ranks, const uint32_t* __restrict__ sizes, uint32_t n, int rb,
                              uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint32_t q = moves[v];
  if (q == NONE) return;
  const uint32_t r = ranks[v];
  keys[2 * (uint64_t)r] = ((uint64_t)part[v] << rb) | r;
  vals[2 * (uint64_t)r] = (uint32_t)(-(int32_t)sizes[v]);
  keys[2 * (uint64_t)r + 1] = ((uint64_t)q << rb) | r;
  vals[2 * (uint64_t)r + 1] = sizes[v];
}
__global__ void k_to_i64(const uint32_t* __restrict__ vals, uint64_t m, int64_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] = (int64_t)(int32_t)vals[i];
}
__global__ void k_size_flags(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                             const int64_t* __restrict__ cum, const uint64_t* __restrict__ poff, uint64_t m, int rb,
                             const uint32_t* __restrict__ psizes, uint32_t omega, int32_t* sv) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t q = (uint32_t)(keys[i] >> rb);
  const uint32_t r = (uint32_t)(keys[i] & (((uint64_t)1 << rb) - 1));
  const uint64_t st = poff[q];
  const int64_t before_seg = st == 0 ? 0 : cum[st - 1];
  const int64_t c = cum[i] - before_seg;
  const int64_t d = (int64_t)(int32_t)vals[i];
  const int64_t cur = (int64_t)psizes[q] + c, prev = cur - d, om = (int64_t)omega;
  const int64_t ex = (cur > om ? cur - om : 0) - (prev > om ? prev - om : 0);
  atomicAdd(&sv[r], (int32_t)ex);
}

// inbound events: (partition, hedge, rank) +-1 for every inbound hedge of every move
// events of partitions in [plo, phi) only (the stage runs in chunks of partitions when all
// events at once would not fit the device)
__global__ void k_inbound_event_counts(const uint32_t* __restrict__ part, const uint32_t* __restrict__ moves,
                                       const uint32_t* __restrict__ inb, uint32_t n, uint32_t plo, uint32_t phi,
                                       uint32_t* __restrict__ cnt) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint32_t q = moves[v], p = part[v];
  const uint32_t in = q == NONE ? 0u : (uint32_t)(p >= plo && p < phi) + (uint32_t)(q >= plo && q < phi);
  cnt[v] = in * inb[v];
}
__global__ void k_inbound_events(const uint32_t* __restrict__ part, const uint32_t* __restrict__ moves,
                                 const uint32_t* __restrict__ ranks, const uint32_t* __restrict__ tch,
                                 const uint64_t* __restrict__ toff, const uint32_t* __restrict__ inb,
                                 const uint64_t* __restrict__ eoff, uint32_t n, int eb, int rb, uint32_t plo,
                                 uint32_t phi, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint32_t q = moves[v];
  if (q == NONE) return;
  const uint32_t p = part[v], r = ranks[v];
  const bool pin = p >= plo && p < phi, qin = q >= plo && q < phi;
  uint64_t o = eoff[v];
  for (uint64_t j = toff[v]; j < toff[v] + inb[v]; ++j) {
    const uint64_t e = tch[j];
    if (pin) {
      keys[o] = ((uint64_t)p << (eb + rb)) | (e << rb) | r;
      vals[o++] = 0xFFFFFFFFu;  // -1
    }
    if (qin) {
      keys[o] = ((uint64_t)q << (eb + rb)) | (e << rb) | r;
      vals[o++] = 1u;
    }
  }
}
// count of destination pins of hedge e in partition q (0 when absent).  The lists start at
// the hedge's pin offset; a list of one entry is read from the record.  Lists of more than
// PP_SMALL entries (only hedges of more than PP_SMALL pins have them) are sorted by
// partition id: binary search; shorter ones are scanned.
__device__ __forceinline__ int64_t dst_count(const uint4* hinfo, const uint64_t* hoff, const uint2* pp,
                                             const uint32_t* pp_dcnt, uint32_t e, uint32_t q) {
  const uint4 hi = hinfo[e];
  const uint64_t s = hoff[e];
  if (hi.z == 1) return hi.x == q ? pp_dcnt[s] : 0;
  if (hi.z <= PP_SMALL) {
    for (uint32_t i = 0; i < hi.z; ++i)
      if (pp[s + i].x == q) return pp_dcnt[s + i];
    return 0;
  }
  uint32_t lo = 0, n = hi.z;  // first entry with partition >= q
  while (n > 0) {
    const uint32_t h = n / 2;
    if (pp[s + lo + h].x < q) {
      lo += h + 1;
      n -= h + 1;
    } else {
      n = h;
    }
  }
  return lo < hi.z && pp[s + lo].x == q ? pp_dcnt[s + lo] : 0;
}
// The first event of every (partition, hedge) run looks up the run's starting destination
// count once (a scan of the hedge's partition list); the run's other events read it there.
__global__ void k_inbound_group_init(const uint64_t* __restrict__ keys, uint64_t m, int eb, int rb,
                                     const uint4* __restrict__ hinfo, const uint64_t* __restrict__ hoff,
                                     const uint2* __restrict__ pp,
                                     const uint32_t* __restrict__ pp_dcnt, int64_t* __restrict__ ginit) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint64_t key = keys[i];
  if (i > 0 && (keys[i - 1] >> rb) == (key >> rb)) return;
  const uint32_t q = (uint32_t)(key >> (eb + rb));
  const uint32_t e = (uint32_t)((key >> rb) & (((uint64_t)1 << eb) - 1));
  ginit[i] = dst_count(hinfo, hoff, pp, pp_dcnt, e, q);
}
__global__ void k_inbound_transitions(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                      const int64_t* __restrict__ cum, uint64_t m, int rb,
                                      const int64_t* __restrict__ ginit, uint32_t* __restrict__ tflag,
                                      int32_t* __restrict__ tdelta) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint64_t key = keys[i];
  const uint64_t grp = key >> rb;  // (partition, hedge)
  // start of this (partition, hedge) run: binary search backwards
  uint64_t lo = 0, hi = i;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if ((keys[mid] >> rb) < grp) lo = mid + 1;
    else hi = mid;
  }
  const int64_t before = lo == 0 ? 0 : cum[lo - 1];
  const int64_t init = ginit[lo];
  const int64_t cur = init + (cum[i] - before);
  const int64_t prev = cur - (int64_t)(int32_t)vals[i];
  int32_t d = 0;
  if (prev == 0 && cur > 0) d = 1;
  else if (prev > 0 && cur == 0) d = -1;
  tflag[i] = d != 0;
  tdelta[i] = d;
}
__global__ void k_inbound_size_events(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ tflag,
                                      const int32_t* __restrict__ tdelta, const uint64_t* __restrict__ tpos, uint64_t m,
                                      int eb, int rb, uint64_t* __restrict__ okeys, uint32_t* __restrict__ ovals) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !tflag[i]) return;
  const uint64_t key = keys[i];
  const uint64_t q = key >> (eb + rb);
  const uint64_t r = key & (((uint64_t)1 << rb) - 1);
  okeys[tpos[i]] = (q << rb) | r;
  ovals[tpos[i]] = (uint32_t)tdelta[i];
}
__global__ void k_inbound_flags(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                const int64_t* __restrict__ cum, const uint64_t* __restrict__ poff, uint64_t m, int rb,
                                const uint32_t* __restrict__ isz, uint32_t delta, int32_t* iv) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t q = (uint32_t)(keys[i] >> rb);
  const uint32_t r = (uint32_t)(keys[i] & (((uint64_t)1 << rb) - 1));
  const uint64_t st = poff[q];
  const int64_t before_seg = st == 0 ? 0 : cum[st - 1];
  const int64_t cur = (int64_t)isz[q] + (cum[i] - before_seg);
  const int64_t prev = cur - (int64_t)(int32_t)vals[i];
  // the first event of a partition assumes it started valid (kernels/refinement.cu:643-644)
  const bool was_valid = i == st || prev <= (int64_t)delta, is_valid = cur <= (int64_t)delta;
  if (was_valid && !is_valid) atomicAdd(&iv[r], 1);
  if (!was_valid && is_valid) atomicSub(&iv[r], 1);
}

// ---- 6. best prefix and apply ---------------------------------------------------------------
__device__ __forceinline__ bool better(const Best& a, const Best& b) {  // a strictly better than b
  if (b.r == NONE) return a.r != NONE;
  if (a.r == NONE) return false;
  if (a.ok != b.ok) return a.ok > b.ok;
  if (!a.ok) return a.r < b.r;
  if (a.sv != b.sv) return a.sv < b.sv;
  if (a.seq != b.seq) return a.seq > b.seq;
  return a.r < b.r;
}
__device__ Best block_best(Best x) {
  __shared__ Best sb[4];
  for (int o = 32; o > 0; o >>= 1) {
    Best y;
    y.ok = __shfl_xor(x.ok, o, 64);
    y.sv = __shfl_xor(x.sv, o, 64);
    y.seq = __shfl_xor(x.seq, o, 64);
    y.r = __shfl_xor(x.r, o, 64);
    y.iv = __shfl_xor(x.iv, o, 64);
    if (better(y, x)) x = y;
  }
  if (lane_id() == 0) sb[threadIdx.x / 64] = x;
  __syncthreads();
  Best b = sb[0];
  for (int i = 1; i < (int)(blockDim.x / 64); ++i)
    if (better(sb[i], b)) b = sb[i];
  return b;
}
__global__ void __launch_bounds__(256) k_best_partial(const int64_t* __restrict__ seq, const int32_t* __restrict__ sv,
                                                      const int32_t* __restrict__ iv, uint32_t V, Best* __restrict__ out) {
  Best x;
  x.r = NONE;
  x.ok = 0;
  x.sv = 0;
  x.seq = 0;
  x.iv = 0;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < V; r += gridDim.x * blockDim.x) {
    Best y;
    y.r = r;
    y.ok = iv[r] == 0;
    y.sv = sv[r];
    y.seq = seq[r];
    y.iv = iv[r];
    if (better(y, x)) x = y;
  }
  Best b = block_best(x);
  if (threadIdx.x == 0) out[blockIdx.x] = b;
}
__global__ void __launch_bounds__(256) k_best_final(const Best* __restrict__ part, uint32_t m, Best* __restrict__ out) {
  Best x;
  x.r = NONE;
  x.ok = 0;
  x.sv = 0;
  x.seq = 0;
  x.iv = 0;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x)
    if (better(part[i], x)) x = part[i];
  Best b = block_best(x);
  if (threadIdx.x == 0) *out = b;
}
// Moves of rank <= best.  With k <= APPLY_LDS_K the size changes are summed per partition in
// LDS and added once per workgroup (with few partitions, one atomic per moved node would
// queue every node of the grid on the same few addresses).
constexpr uint32_t APPLY_LDS_K = 1024;
__global__ void __launch_bounds__(256) k_apply(const uint32_t* __restrict__ moves, const uint32_t* __restrict__ ranks,
                                               const uint32_t* __restrict__ sizes, uint32_t n, uint32_t best,
                                               uint32_t* __restrict__ part, uint32_t* psizes, uint32_t k) {
  __shared__ long long s_d[APPLY_LDS_K];
  const bool lds = k <= APPLY_LDS_K;
  if (lds) {
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) s_d[i] = 0;
    __syncthreads();
  }
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) {
    const uint32_t q = moves[v];
    if (q != NONE && ranks[v] <= best) {
      const uint32_t s = sizes[v], p = part[v];
      if (lds) {
        atomicAdd((unsigned long long*)&s_d[p], (unsigned long long)(-(long long)s));
        atomicAdd((unsigned long long*)&s_d[q], (unsigned long long)s);
      } else {
        atomicSub(&psizes[p], s);
        atomicAdd(&psizes[q], s);
      }
      part[v] = q;
    }
  }
  if (lds) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x)
      if (s_d[i]) atomicAdd(&psizes[i], (uint32_t)s_d[i]);  // two's complement: a net decrease wraps back
  }
}

__global__ void k_connectivity(const uint32_t* __restrict__ lambda, const float* __restrict__ w, uint32_t E,
                               unsigned long long* out) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  int64_t c = 0;
  if (e < E && lambda[e] > 0) c = (int64_t)fixed_weight(w[e]) * (int64_t)(lambda[e] - 1);
  c = wave_sum(c);
  if (lane_id() == 0 && c) atomicAdd(out, (unsigned long long)c);
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx, uint32_t m,
                             uint32_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] = src[idx[i]];
}

__global__ void k_max_u32(const uint32_t* __restrict__ p, uint32_t n, uint32_t* out) {
  uint32_t m = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = p[i] > m ? p[i] : m;
  m = wave_max(m);
  if (lane_id() == 0) atomicMax(out, m);
}

}  // namespace

// ==========================================================================================
// host side

namespace {

struct PartitionLists {
  DBuf<uint2> pp;      // (partition, pin count) per list entry, at the hedge's own offsets
  DBuf<uint4> hinfo;   // per hedge: list start (lo, hi) or its one entry, length, fixed-point weight
  DBuf<uint32_t> dcnt, lambda, inbsz;
  DBuf<uint4> hm;      // instead of the lists (k-way, k <= MASK_K): per hedge, its partition masks
  uint64_t P = 0;      // pins of the level
};

// Partition lists of every hedge.  want_inbound adds the destination counts and the
// per-partition inbound sizes (the inbound-set constraint needs them; k-way does not).
// The hedge classes of a level depend only on hedge lengths and k, so refine() computes them
// in its first round and reuses them: for the partition-list kernels the pin count, the
// hedges of more than PP_SMALL pins with LDS tables, those with HBM tables and their table
// offsets; for the cascade the three classes of hedges above CAS_SMALL pins (for the
// sort_max they were built with).
struct LevelLayout {
  bool ready = false;
  uint64_t P = 0;
  uint32_t nl = 0, ng = 0;
  DBuf<uint32_t> large, glist;
  DBuf<uint64_t> goff;
  uint64_t gslots = 0;
  uint32_t cas_sort_max = 0xFFFFFFFFu;
  DBuf<uint32_t> cas_list[3];
  uint32_t cas_n[3] = {0, 0, 0};
  // hedges of 2..CAS_SMALL pins as their own CSR, in order of their first pin
  SmallCsr cs;
};

// masks: only the partition masks (k-way with k <= MASK_K: nothing else reads the lists).
void build_partition_lists(Ctx& c, const RefineLevel& L, const uint32_t* part, PartitionLists& pl, bool want_inbound,
                           LevelLayout* layout = nullptr, bool masks = false) {
  LevelLayout local;
  LevelLayout& Y = layout ? *layout : local;
  if (!Y.ready) Y.P = L.e ? read_scalar(c, L.hoff + L.e) : 0;
  const uint64_t P = Y.P;
  pl.P = P;
  if (masks) {
    DBuf<uint32_t> lflag;
    if (!Y.ready) lflag = DBuf<uint32_t>(c, L.e);
    pl.hm = DBuf<uint4>(c, std::max<uint32_t>(L.e, 1));
    HGP_LAUNCH(k_hmask_small, ceil_div(L.e, 256), 256, 0, c, L.hedges, L.hoff, part, L.e, pl.hm.get(), lflag.get());
    if (!Y.ready) {  // the hedges above PP_SMALL pins (no HBM tables at k <= MASK_K)
      Y.ready = true;
      DBuf<uint64_t> pos(c, (uint64_t)L.e + 1);
      exclusive_scan(c, lflag.get(), pos.get(), L.e);
      Y.nl = (uint32_t)read_scalar(c, pos.get() + L.e);
      if (Y.nl) {
        Y.large = DBuf<uint32_t>(c, Y.nl);
        DBuf<uint32_t> small(c, L.e - Y.nl);
        HGP_LAUNCH(k_split, ceil_div(L.e, 256), 256, 0, c, lflag.get(), pos.get(), L.e, Y.large.get(), small.get());
      }
    }
    if (Y.nl) HGP_LAUNCH(k_hmask_large, Y.nl, 64, 0, c, L.hedges, L.hoff, part, Y.large.get(), pl.hm.get());
    return;
  }
  pl.pp = DBuf<uint2>(c, P);
  pl.hinfo = DBuf<uint4>(c, L.e);
  pl.lambda = DBuf<uint32_t>(c, L.e);
  if (want_inbound) {
    pl.dcnt = DBuf<uint32_t>(c, P);
    pl.inbsz = DBuf<uint32_t>(c, std::max<uint32_t>(L.k, 1));
    memset_async(c, pl.inbsz.get(), 0, (uint64_t)L.k * sizeof(uint32_t));
  }
  DBuf<uint32_t> lflag;
  if (!Y.ready) lflag = DBuf<uint32_t>(c, L.e);
  if (want_inbound) {
    const uint32_t k_lds = L.k <= PP_HIST_CAP ? L.k : 0u;
    HGP_LAUNCH((k_pp_small<true>), ceil_div(L.e, 256), 256, k_lds * sizeof(uint32_t), c, L.hedges, L.hoff, L.srcc, part,
               L.w, L.e, pl.pp.get(), pl.dcnt.get(), pl.lambda.get(), pl.hinfo.get(), lflag.get(), pl.inbsz.get(),
               k_lds);
  } else {
    HGP_LAUNCH((k_pp_small<false>), ceil_div(L.e, 256), 256, 0, c, L.hedges, L.hoff, L.srcc, part, L.w, L.e,
               pl.pp.get(), (uint32_t*)nullptr, pl.lambda.get(), pl.hinfo.get(), lflag.get(), (uint32_t*)nullptr, 0u);
  }
  // Distinct partitions per hedge <= min(k, len).  Tables of up to PP_LDS_CAP slots sit in
  // LDS; hedges needing more (len and k both above ~2K) get HBM tables.
  const uint32_t lds_max_len = (PP_LDS_CAP - 64) / 2;
  const uint32_t cap = pp_table_cap(lds_max_len, L.k);  // <= PP_LDS_CAP
  if (!Y.ready) {
    Y.ready = true;
    DBuf<uint64_t> pos(c, (uint64_t)L.e + 1);
    exclusive_scan(c, lflag.get(), pos.get(), L.e);
    Y.nl = (uint32_t)read_scalar(c, pos.get() + L.e);
    if (Y.nl == 0) return;
    Y.large = DBuf<uint32_t>(c, Y.nl);
    DBuf<uint32_t> small(c, L.e - Y.nl);
    // k_split puts flagged (large) entries first list
    HGP_LAUNCH(k_split, ceil_div(L.e, 256), 256, 0, c, lflag.get(), pos.get(), L.e, Y.large.get(), small.get());
    if (L.k > lds_max_len) {
      DBuf<uint32_t> gflag(c, Y.nl);
      DBuf<uint64_t> gpos(c, (uint64_t)Y.nl + 1);
      HGP_LAUNCH(k_flag_longer, ceil_div(Y.nl, 256), 256, 0, c, Y.large.get(), Y.nl, L.hoff, (uint64_t)lds_max_len,
                 gflag.get());
      exclusive_scan(c, gflag.get(), gpos.get(), Y.nl);
      Y.ng = (uint32_t)read_scalar(c, gpos.get() + Y.nl);
      if (Y.ng) {
        DBuf<uint32_t> lds_list(c, Y.nl - Y.ng);
        Y.glist = DBuf<uint32_t>(c, Y.ng);
        HGP_LAUNCH(k_split_list, ceil_div(Y.nl, 256), 256, 0, c, gflag.get(), gpos.get(), Y.large.get(), Y.nl,
                   Y.glist.get(), lds_list.get());
        Y.large = std::move(lds_list);
        Y.nl -= Y.ng;
        DBuf<uint64_t> gcap(c, Y.ng);
        Y.goff = DBuf<uint64_t>(c, (uint64_t)Y.ng + 1);
        HGP_LAUNCH(k_pp_gcap, ceil_div(Y.ng, 256), 256, 0, c, Y.glist.get(), Y.ng, L.hoff, L.k, gcap.get());
        exclusive_scan(c, gcap.get(), Y.goff.get(), Y.ng);
        Y.gslots = read_scalar(c, Y.goff.get() + Y.ng);
      }
    }
  }
  auto launch = [&](auto inb, auto gmem, const uint32_t* lst, uint32_t m, const uint64_t* goff, uint32_t* gtab) {
    constexpr bool INB = decltype(inb)::value, GMEM = decltype(gmem)::value;
    if (m == 0) return;
    HGP_LAUNCH((k_pp_large<INB, GMEM>), m, 64, GMEM ? 0 : 3 * cap * sizeof(uint32_t), c, L.hedges, L.hoff, L.srcc, part,
               L.w, lst, cap, L.k, goff, gtab, pl.pp.get(), INB ? pl.dcnt.get() : (uint32_t*)nullptr, pl.lambda.get(),
               pl.hinfo.get(), INB ? pl.inbsz.get() : (uint32_t*)nullptr);
  };
  using T = std::true_type;
  using F = std::false_type;
  if (want_inbound) launch(T{}, F{}, Y.large.get(), Y.nl, nullptr, nullptr);
  else launch(F{}, F{}, Y.large.get(), Y.nl, nullptr, nullptr);
  if (Y.ng) {
    DBuf<uint32_t> gtab(c, 3 * Y.gslots);
    if (want_inbound) launch(T{}, T{}, Y.glist.get(), Y.ng, Y.goff.get(), gtab.get());
    else launch(F{}, T{}, Y.glist.get(), Y.ng, Y.goff.get(), gtab.get());
  }
}

uint32_t discount_of(uint32_t rep, uint32_t R) { return rep < R / 3 ? 1u : (rep < 2 * R / 3 ? 2u : NONE); }

// peak device bytes per inbound-set event: keys and values with their sort buffers and tile
// counts, running sums, transition flags, deltas, group starts and positions
constexpr uint64_t INBOUND_EVENT_BYTES = 64;

uint64_t event_budget(Ctx& c) {
  if (c.event_budget) return c.event_budget;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    return 16ull << 30;
  }
  return std::max<uint64_t>(1ull << 26, ((uint64_t)free_b + c.cached) / 2);
}


// Algorithmic bytes of the gain kernels over `list` (evaluated only while the kernel is
// profiled): per node 36 (partition, size, two offsets, move and gain out) + 20 per touching
// hedge (hedge id and its 16-byte record: list start, list length, weight) + 8 per listed
// partition entry (partition and pin count).
void compute_gains(Ctx& c, const RefineLevel& L, const PartitionLists& pl, int mode, uint32_t omega, uint32_t repeat,
                   uint32_t discount, const uint32_t* part, const uint32_t* psizes, uint32_t* moves, int64_t* gains) {
  // hash order of partitions for the non-adjacent fallback, sorted on the device
  const uint32_t kk = std::max<uint32_t>(L.k, 1);
  DBuf<uint32_t> okeys(c, kk), dorder(c, kk);
  HGP_LAUNCH(k_order_keys, ceil_div(L.k, 256), 256, 0, c, L.k, repeat, okeys.get(), dorder.get());
  radix_sort(c, okeys.get(), dorder.get(), L.k, 32);
  GainArgs a{};
  a.tch = L.tch;
  a.toff = L.toff;
  a.hinfo = pl.hinfo.get();
  a.pp = pl.pp.get();
  a.part = part;
  a.sizes = L.sizes;
  a.psizes = psizes;
  a.order = dorder.get();
  a.k = L.k;
  a.omega = omega;
  a.discount = discount;
  a.repeat = repeat;
  a.incc = mode == HGP_MODE_INCC;
  a.moves = moves;
  a.gains = gains;
  if (pl.hm.get()) {  // k-way from partition masks: tiny nodes in the router, the rest a wave each
    // the sub-lists start at their XCD's first node: room for the route grid's 256 nodes per workgroup
    const uint32_t rg = std::max<uint32_t>(1, ceil_div(L.n, 256));
    DBuf<uint32_t> ml(c, (uint64_t)rg * 256), mc(c, XCDS);
    memset_async(c, mc.get(), 0, 4 * XCDS);
    HGP_LAUNCH(k_gains_route_mask, rg, 256, 0, c, a, pl.hm.get(), L.w, L.n, ml.get(), mc.get());
    a.prof_bytes = prof_match(c, "k_gains_mask") && c.prof_count ? c.prof_dev : nullptr;
    // waves per node from the level's mean degree: a wave per node while the nodes are many
    // and light, 4 or 16 when a wave would walk thousands of hedges alone
    const uint64_t mean = L.n ? pl.P / L.n : 0;
    const uint32_t W = mean >= c.tune.mask_wide16_degree ? 16 : (mean >= c.tune.mask_wide4_degree ? 4 : 1);
    // every XCD needs a workgroup for its sub-list: grids of at least XCDS
    const uint32_t wg = std::max<uint32_t>(XCDS, std::min<uint32_t>(L.n, LIGHT_GRID));
    const bool sk = L.k <= MASK_SMALL_K;
    if (W == 16 && sk)
      HGP_LAUNCH((k_gains_mask_wide<16, true>), wg, 1024, 0, c, a, pl.hm.get(), L.w, ml.get(), mc.get(), rg);
    else if (W == 16)
      HGP_LAUNCH((k_gains_mask_wide<16, false>), wg, 1024, 0, c, a, pl.hm.get(), L.w, ml.get(), mc.get(), rg);
    else if (W == 4 && sk)
      HGP_LAUNCH((k_gains_mask_wide<4, true>), wg, 256, 0, c, a, pl.hm.get(), L.w, ml.get(), mc.get(), rg);
    else if (W == 4)
      HGP_LAUNCH((k_gains_mask_wide<4, false>), wg, 256, 0, c, a, pl.hm.get(), L.w, ml.get(), mc.get(), rg);
    else
      HGP_LAUNCH(k_gains_mask, std::max<uint32_t>(XCDS, std::min<uint32_t>(ceil_div(L.n, 4), LIGHT_GRID)), 256, 0, c,
                 a, pl.hm.get(), L.w, ml.get(), mc.get(), rg);
    return;
  }
  // tiny nodes are scored by the router; the light and block lists and their counts stay on
  // the device, and the persistent light and block kernels stride over them
  DBuf<uint64_t> bound(c, L.n);
  DBuf<uint32_t> ll(c, L.n), bl(c, L.n), counts(c, 2);
  memset_async(c, counts.get(), 0, 8);
  HGP_LAUNCH(k_gains_route, ceil_div(L.n, 256), 256, 0, c, a, L.n, bound.get(), ll.get(), bl.get(), counts.get());
  const uint32_t capmax = L.k <= GAIN_CAP ? std::max<uint32_t>(L.k, 1) : GAIN_CAP;
  const uint32_t lgrid = std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(L.n, 4), LIGHT_GRID));
  // when this kernel is profiled it counts its algorithmic bytes itself (no extra pass)
  a.prof_bytes = prof_match(c, "k_gains_light") && c.prof_count ? c.prof_dev : nullptr;
  HGP_LAUNCH(k_gains_light, lgrid, 256, 4 * capmax * 12, c, a, ll.get(), counts.get(), bound.get(), capmax);
  a.prof_bytes = nullptr;
  DBuf<uint32_t> hl;
  uint32_t nh = 0;
  // workgroup per node with an LDS table; hashed tables that would overflow go to HBM
  const uint32_t bgrid = std::max<uint32_t>(1, std::min<uint32_t>(L.n, BLOCK_GRID));
  if (L.k <= GB_DIRECT) {
    HGP_LAUNCH((k_gains_block<true>), bgrid, 1024, 0, c, a, bl.get(), counts.get() + 1, (uint32_t*)nullptr,
               (uint32_t*)nullptr);
    return;
  }
  DBuf<uint32_t> ovf(c, L.n), novf(c, 1);
  memset_async(c, novf.get(), 0, 4);
  HGP_LAUNCH((k_gains_block<false>), bgrid, 1024, 0, c, a, bl.get(), counts.get() + 1, ovf.get(), novf.get());
  nh = (uint32_t)read_scalar(c, novf.get());
  if (nh == 0) return;
  hl = std::move(ovf);
  DBuf<uint64_t> caps(c, nh);
  HGP_LAUNCH(k_heavy_tabcap, ceil_div(nh, 256), 256, 0, c, hl.get(), bound.get(), nh, L.k, caps.get());
  std::vector<uint64_t> hc(nh);
  copy_d2h(c, hc.data(), caps.get(), (uint64_t)nh * sizeof(uint64_t));
  uint32_t b0 = 0;
  while (b0 < nh) {
    uint64_t tot = 0;
    uint32_t b1 = b0;
    while (b1 < nh && (b1 == b0 || tot + hc[b1] <= HEAVY_BATCH_SLOTS)) tot += hc[b1++];
    std::vector<uint64_t> off(b1 - b0 + 1, 0);
    for (uint32_t i = b0; i < b1; ++i) off[i - b0 + 1] = off[i - b0] + hc[i];
    DBuf<uint64_t> doff(c, off.size());
    copy_h2d(c, doff.get(), off.data(), off.size() * sizeof(uint64_t));
    DBuf<uint32_t> gk(c, tot);
    DBuf<int64_t> gc(c, tot);
    HGP_LAUNCH(k_gains_heavy, b1 - b0, 256, 0, c, a, hl.get() + b0, doff.get(), gk.get(), gc.get());
    sync(c);
    b0 = b1;
  }
}

void rank_by_gain(Ctx& c, uint32_t n, const uint32_t* moves, const int64_t* gains, uint32_t* ranks) {
  DBuf<uint64_t> keys(c, n);
  DBuf<uint32_t> vals(c, n), k2(c, n);
  HGP_LAUNCH(k_rank_keys_gain, ceil_div(n, 256), 256, 0, c, moves, gains, n, keys.get(), vals.get());
  radix_sort_range_buf(c, keys, vals, n);
  HGP_LAUNCH(k_invalid_keys, ceil_div(n, 256), 256, 0, c, moves, vals.get(), n, k2.get());
  radix_sort_buf(c, k2, vals, n, 0, 1);
  HGP_LAUNCH(k_scatter_ranks, ceil_div(n, 256), 256, 0, c, vals.get(), n, ranks);
}

void rank_by_chaining(Ctx& c, uint32_t n, uint32_t k, const uint32_t* part, const uint32_t* moves, const int64_t* gains,
                      const uint32_t* sizes, uint32_t* ranks) {
  const uint32_t g = ceil_div(n, 256);
  // sorted space: stable by (src partition asc, valid first, gain desc), ties by node id
  DBuf<uint64_t> k1(c, n);
  DBuf<uint32_t> orig(c, n), k2(c, n);
  HGP_LAUNCH(k_chain_sortkeys, g, 256, 0, c, part, moves, gains, n, k1.get(), orig.get());
  radix_sort_range_buf(c, k1, orig, n);
  HGP_LAUNCH(k_chain_sortkeys2, g, 256, 0, c, part, moves, orig.get(), n, k2.get());
  radix_sort_buf(c, k2, orig, n, 0, bit_width(2ull * k + 1));
  k1.reset();
  DBuf<uint32_t> src(c, n), dst(c, n), sz(c, n), ok(c, n), ob(c, n), oe(c, n), next(c, n), prev(c, n);
  DBuf<int64_t> w(c, n);
  DBuf<unsigned long long> claim(c, n);
  DBuf<uint32_t> progress(c, 4);
  DBuf<uint4> pk(c, n);
  DBuf<uint32_t> from(c, n);
  fill_u32(c, progress.get(), 0u, 4);
  ChainArrays ca{orig.get(), src.get(), dst.get(), sz.get(), w.get(), ok.get(), ob.get(), oe.get(),
                 next.get(), prev.get(), claim.get(), progress.get(), pk.get(), from.get()};
  HGP_LAUNCH(k_chain_gather, g, 256, 0, c, ca, part, moves, gains, sizes, n);
  DBuf<uint64_t> poff(c, (uint64_t)k + 1);
  // src partitions of sorted space are the k2 keys >> 1
  segment_offsets(c, k2.get(), 1, n, k, poff.get());
  HGP_LAUNCH(k_chain_ranges, g, 256, 0, c, ca, poff.get(), n);
  for (int it = 0; it < 4; ++it) {
    HGP_LAUNCH(k_chain_propose, ceil_div(n, RESOLVE_BLOCK), RESOLVE_BLOCK, 0, c, ca, n, it);
    HGP_LAUNCH(k_chain_commit, g, 256, 0, c, ca, n, it);
  }
  DBuf<uint32_t> comp(c, n), pos(c, n), cflag(c, n), cinv(c, n), csize(c, n), odd(c, 1);
  DBuf<long long> cw(c, n);
  HGP_LAUNCH(k_zero3, g, 256, 0, c, (unsigned long long*)cw.get(), cflag.get(), cinv.get(), (uint64_t)n);
  fill_u32(c, csize.get(), 0u, n);
  fill_u32(c, odd.get(), 0u, 1);
  HGP_LAUNCH(k_chain_comp, g, 256, 0, c, ca, n, comp.get(), pos.get(), cflag.get(), cw.get(), cinv.get(), csize.get(),
             odd.get());
  DBuf<uint64_t> cpos(c, (uint64_t)n + 1), mo(c, 2);
  exclusive_scan(c, cflag.get(), cpos.get(), n);
  HGP_LAUNCH(k_pack2, 1, 1, 0, c, cpos.get() + n, odd.get(), mo.get());  // one read-back for both
  uint64_t mo_h[2];
  copy_d2h(c, mo_h, mo.get(), sizeof mo_h);
  const uint32_t m = (uint32_t)mo_h[0];
  const bool direct = mo_h[1] == 0;
  DBuf<uint32_t> clist(c, m), cinvk(c, m), crank(c, n);
  DBuf<uint64_t> ckey(c, m);
  HGP_LAUNCH(k_comp_list, g, 256, 0, c, cflag.get(), cpos.get(), n, cw.get(), cinv.get(), clist.get(), ckey.get(),
             cinvk.get());
  // components by (valid first, summed gain desc, head id asc): the list starts in ascending
  // head id; a stable sort on the summed gain, then a stable 1-bit pass that moves the
  // components made of an invalid move (summed gain 0) behind every valid one
  radix_sort_range_buf(c, ckey, clist, m);
  DBuf<uint32_t> inv2(c, m);
  HGP_LAUNCH(k_gather_u32, ceil_div(m, 256), 256, 0, c, cinv.get(), clist.get(), m, inv2.get());
  radix_sort_buf(c, inv2, clist, m, 0, 1);
  HGP_LAUNCH(k_comp_rank, ceil_div(m, 256), 256, 0, c, clist.get(), m, crank.get());
  if (direct) {  // every chain has its head: ranks from component starts and distances
    DBuf<uint32_t> rsz(c, m);
    DBuf<uint64_t> start(c, (uint64_t)m + 1);
    HGP_LAUNCH(k_comp_sizes, ceil_div(m, 256), 256, 0, c, clist.get(), csize.get(), m, rsz.get());
    exclusive_scan(c, rsz.get(), start.get(), m);
    HGP_LAUNCH(k_chain_direct_ranks, g, 256, 0, c, comp.get(), pos.get(), crank.get(), start.get(), orig.get(), n,
               ranks);
    return;
  }
  // final order: (component rank, position), ties in sorted-space index order: the keys are
  // laid out in index order and the sort is stable
  DBuf<uint64_t> ek(c, n);
  DBuf<uint32_t> ev(c, n);
  HGP_LAUNCH(k_edge_keys, g, 256, 0, c, comp.get(), pos.get(), crank.get(), n, ek.get(), ev.get());
  radix_sort_buf(c, ek, ev, n, 0, bit_width(m) + 8);
  HGP_LAUNCH(k_edge_ranks, g, 256, 0, c, ev.get(), orig.get(), n, ranks);
}

struct RoundState {
  DBuf<uint32_t> moves, ranks;
  DBuf<int64_t> gains;
  DBuf<unsigned long long> seq;  // int64 values, indexed by rank
  DBuf<int32_t> sv, iv;
  uint32_t V = 0;
  Best best;
};

void analyze_round(Ctx& c, const RefineLevel& L, int mode, uint32_t omega, uint32_t delta, uint32_t R, uint32_t rep,
                   bool chainup, const uint32_t* part, const uint32_t* psizes, RoundState& st,
                   LevelLayout* layout = nullptr) {
  const uint32_t n = L.n;
  const uint32_t g = ceil_div(n, 256);
  const bool incc = mode == HGP_MODE_INCC;
  PartitionLists pl;
  build_partition_lists(c, L, part, pl, incc, layout, !incc && L.k <= MASK_K);
  st.moves = DBuf<uint32_t>(c, n);
  st.gains = DBuf<int64_t>(c, n);
  st.ranks = DBuf<uint32_t>(c, n);
  compute_gains(c, L, pl, mode, omega, rep, discount_of(rep, R), part, psizes, st.moves.get(), st.gains.get());
  if (chainup) rank_by_gain(c, n, st.moves.get(), st.gains.get(), st.ranks.get());
  else rank_by_chaining(c, n, L.k, part, st.moves.get(), st.gains.get(), L.sizes, st.ranks.get());
  {
    DBuf<uint32_t> cnt(c, 1);
    memset_async(c, cnt.get(), 0, 4);
    HGP_LAUNCH(k_count_valid, std::min<uint32_t>(g, 1024), 256, 0, c, st.moves.get(), n, cnt.get());
    st.V = read_scalar(c, cnt.get());
  }
  const uint32_t V = st.V;
  st.seq = DBuf<unsigned long long>(c, std::max<uint32_t>(V, 1));
  st.sv = DBuf<int32_t>(c, std::max<uint32_t>(V, 1));
  st.iv = DBuf<int32_t>(c, std::max<uint32_t>(V, 1));
  st.best.r = NONE;
  if (V == 0) return;
  HGP_LAUNCH(k_zero3, ceil_div(V, 256), 256, 0, c, st.seq.get(), (uint32_t*)st.sv.get(), (uint32_t*)st.iv.get(),
             (uint64_t)V);
  // cascade.  The hedges of 2..CAS_SMALL pins first, alone on the stream; then, with bit 1 of
  // the side_streams tuning, the sorted and direct classes on side streams beside the size and
  // inbound events (all classes add their terms to seq with atomics; the node records stay
  // alive until the join before seq's prefix sum).
  DBuf<uint4> pm;
  DBuf<uint2> pm2;
  std::optional<Fork> cas_fork;
  {
    const bool packed = L.k < PM2_MAX_K;
    if (packed) {
      pm2 = DBuf<uint2>(c, n);
      HGP_LAUNCH(k_pack_nodes2, g, 256, 0, c, part, st.moves.get(), st.ranks.get(), n, pm2.get());
    } else {
      pm = DBuf<uint4>(c, n);
      HGP_LAUNCH(k_pack_nodes, g, 256, 0, c, part, st.moves.get(), st.ranks.get(), n, pm.get());
    }
    CascadeArgs a{L.hedges, L.hoff, L.w, pm.get(), pm2.get(), incc ? 1 : 0, st.seq.get()};
    const int crb = std::max(1, bit_width(V));  // rank + 1 <= V
    const int cpb = std::max(1, bit_width(L.k));
    const uint32_t sort_max = cpb + crb + 11 < 64 ? CAS_SORT : 0u;
    LevelLayout local;
    LevelLayout& Y = layout ? *layout : local;
    if (Y.cas_sort_max != sort_max) {  // the level's cascade classes (its first round)
      Y.cas_sort_max = sort_max;
      DBuf<uint32_t> fl[4];
      for (int cls = 0; cls < 4; ++cls) fl[cls] = DBuf<uint32_t>(c, L.e);
      DBuf<uint64_t> ks(c, std::max<uint32_t>(L.e, 1));
      HGP_LAUNCH(k_cascade_classes, ceil_div(L.e, 256), 256, 0, c, L.hedges, L.hoff, L.e, sort_max, fl[3].get(),
                 ks.get(), fl[0].get(), fl[1].get(), fl[2].get());
      DBuf<uint64_t> pos(c, (uint64_t)L.e + 1);
      for (int cls = 0; cls < 3; ++cls) {
        exclusive_scan(c, fl[cls].get(), pos.get(), L.e);
        Y.cas_n[cls] = (uint32_t)read_scalar(c, pos.get() + L.e);
        Y.cas_list[cls] = DBuf<uint32_t>(c, Y.cas_n[cls]);
        if (Y.cas_n[cls])
          HGP_LAUNCH(k_compact, ceil_div(L.e, 256), 256, 0, c, fl[cls].get(), pos.get(), L.e, Y.cas_list[cls].get());
      }
      // hedges of 2..CAS_SMALL pins as their own CSR, in order of their first pin
      exclusive_scan(c, fl[3].get(), pos.get(), L.e);
      const uint32_t m = (uint32_t)read_scalar(c, pos.get() + L.e);
      Y.cs = SmallCsr();
      if (m) {
        DBuf<uint64_t> kc(c, m);
        HGP_LAUNCH(k_compact_u64, ceil_div(L.e, 256), 256, 0, c, fl[3].get(), pos.get(), L.e, ks.get(), kc.get());
        radix_sort(c, kc.get(), nullptr, m, 32 + std::max(1, bit_width(L.n)));
        Y.cs = small_csr(c, kc.get(), m, L.hoff, L.hedges, L.w);
      }
    }
    auto small_cascade = [&] {
      const uint32_t m = Y.cs.m;
      if (!m) return;
      const SmallHedges sh{Y.cs.rec.get(), Y.cs.pins.get(), Y.cs.w.get(), m};
      const uint32_t cb = c.tune.cascade_block;  // 256 unless the autotuner chose otherwise
      if (a.pm2) HGP_LAUNCH(k_cascade_small<true>, ceil_div(m, cb), cb, 0, c, a, sh);
      else HGP_LAUNCH(k_cascade_small<false>, ceil_div(m, cb), cb, 0, c, a, sh);
      if (prof_match(c, "k_cascade_small") && c.prof_count && c.prof_dev)
        HGP_LAUNCH(k_cascade_small_bytes, ceil_div(m, 256), 256, 0, c, a, sh, c.prof_dev);
    };
    // bit 2 of side_streams: the side streams start before the small-hedge cascade and run
    // beside it too (otherwise after it, beside the events only)
    const bool early = (c.tune.side_streams & 6u) == 6u;
    if (!early) small_cascade();
    cas_fork.emplace(c, (c.tune.side_streams & 2u) != 0);
    for (int cls = 0; cls < 3; ++cls) {
      const uint32_t m = Y.cas_n[cls];
      if (m == 0) continue;
      const uint32_t* lst = Y.cas_list[cls].get();
      cas_fork->on(cls, [&] {
        if (cls == 0) HGP_LAUNCH((k_cascade_sorted<CAS_SORT_S, 2 * CAS_SORT_S, 64>), m, 64, 0, c, a, lst, crb);
        else if (cls == 1)
          HGP_LAUNCH((k_cascade_sorted<CAS_SORT, CAS_EV, CAS_SORT_NT>), m, CAS_SORT_NT, 0, c, a, lst, crb);
        else HGP_LAUNCH(k_cascade_large, m, 64, 0, c, a, lst);
      });
    }
    if (early) small_cascade();
  }
  const int rb = std::max(1, bit_width(V - 1));
  const int kb = std::max(1, bit_width(L.k));
  // size events
  {
    const uint64_t m = 2ull * V;
    DBuf<uint64_t> keys(c, m), poff(c, (uint64_t)L.k + 1);
    DBuf<uint32_t> vals(c, m);
    DBuf<int64_t> cum(c, m);
    HGP_LAUNCH(k_size_events, g, 256, 0, c, part, st.moves.get(), st.ranks.get(), L.sizes, n, rb, keys.get(), vals.get());
    // the events sit at 2 * rank (+1), so they are already in rank order: sorting by the
    // partition bits alone orders them by (partition, rank)
    radix_sort_buf(c, keys, vals, m, rb, rb + kb);
    HGP_LAUNCH(k_to_i64, ceil_div(m, 256), 256, 0, c, vals.get(), m, cum.get());
    inclusive_scan(c, cum.get(), cum.get(), m);
    segment_offsets(c, keys.get(), rb, m, L.k, poff.get());
    HGP_LAUNCH(k_size_flags, ceil_div(m, 256), 256, 0, c, keys.get(), vals.get(), cum.get(), poff.get(), m, rb, psizes,
               omega, st.sv.get());
    inclusive_scan(c, st.sv.get(), st.sv.get(), V);
  }
  // inbound-set events (INCC only: with k-way's Delta = INT32_MAX every prefix is valid)
  if (incc) {
    const int eb = std::max(1, bit_width(L.e));
    HGP_REQUIRE(kb + eb + rb <= 64, HGP_E_CAPACITY, "inbound events: key exceeds 64 bits");
    DBuf<uint32_t> ecnt(c, n);
    DBuf<uint64_t> eoff(c, (uint64_t)n + 1);
    // All events at once need about INBOUND_EVENT_BYTES each (keys and values with their
    // sort buffers, running sums, transition flags, deltas, group starts, positions); when that
    // exceeds the event budget the partitions are split into ranges whose events are processed
    // one range after the other.  Every event stays with its partition, so the violations
    // summed over the ranges are exact.
    HGP_LAUNCH(k_inbound_event_counts, g, 256, 0, c, part, st.moves.get(), L.inb, n, 0u, L.k, ecnt.get());
    exclusive_scan(c, ecnt.get(), eoff.get(), n);
    const uint64_t m_all = read_scalar(c, eoff.get() + n);
    const uint64_t budget = event_budget(c);
    const uint32_t chunks = (uint32_t)std::min<uint64_t>(
        std::max<uint32_t>(L.k, 1), std::max<uint64_t>(1, (m_all * INBOUND_EVENT_BYTES + budget - 1) / budget));
    if (c.verbose >= 2 && chunks > 1)
      fprintf(stderr, "  inbound events: %llu in %u partition ranges\n", (unsigned long long)m_all, chunks);
    for (uint32_t ch = 0; ch < chunks && m_all; ++ch) {
      const uint32_t plo = (uint32_t)((uint64_t)L.k * ch / chunks), phi = (uint32_t)((uint64_t)L.k * (ch + 1) / chunks);
      if (chunks > 1) {
        HGP_LAUNCH(k_inbound_event_counts, g, 256, 0, c, part, st.moves.get(), L.inb, n, plo, phi, ecnt.get());
        exclusive_scan(c, ecnt.get(), eoff.get(), n);
      }
      const uint64_t m = chunks > 1 ? read_scalar(c, eoff.get() + n) : m_all;
      if (!m) continue;
      DBuf<uint64_t> keys(c, m);
      DBuf<uint32_t> vals(c, m);
      HGP_LAUNCH(k_inbound_events, g, 256, 0, c, part, st.moves.get(), st.ranks.get(), L.tch, L.toff, L.inb, eoff.get(), n,
                 eb, rb, plo, phi, keys.get(), vals.get());
      radix_sort(c, keys.get(), vals.get(), m, kb + eb + rb);
      DBuf<int64_t> cum(c, m);
      HGP_LAUNCH(k_to_i64, ceil_div(m, 256), 256, 0, c, vals.get(), m, cum.get());
      inclusive_scan(c, cum.get(), cum.get(), m);
      DBuf<uint32_t> tflag(c, m);
      DBuf<int32_t> tdelta(c, m);
      DBuf<int64_t> ginit(c, m);
      HGP_LAUNCH(k_inbound_group_init, ceil_div(m, 256), 256, 0, c, keys.get(), m, eb, rb, pl.hinfo.get(), L.hoff,
                 pl.pp.get(), pl.dcnt.get(), ginit.get());
      HGP_LAUNCH(k_inbound_transitions, ceil_div(m, 256), 256, 0, c, keys.get(), vals.get(), cum.get(), m, rb,
                 ginit.get(), tflag.get(), tdelta.get());
      DBuf<uint64_t> tpos(c, m + 1);
      exclusive_scan(c, tflag.get(), tpos.get(), m);
      const uint64_t T = read_scalar(c, tpos.get() + m);
      if (T) {
        DBuf<uint64_t> ok(c, T), poff(c, (uint64_t)L.k + 1);
        DBuf<uint32_t> ov(c, T);
        HGP_LAUNCH(k_inbound_size_events, ceil_div(m, 256), 256, 0, c, keys.get(), tflag.get(), tdelta.get(), tpos.get(),
                   m, eb, rb, ok.get(), ov.get());
        keys.reset();
        radix_sort(c, ok.get(), ov.get(), T, rb + kb);
        DBuf<int64_t> cum2(c, T);
        HGP_LAUNCH(k_to_i64, ceil_div(T, 256), 256, 0, c, ov.get(), T, cum2.get());
        inclusive_scan(c, cum2.get(), cum2.get(), T);
        segment_offsets(c, ok.get(), rb, T, L.k, poff.get());
        HGP_LAUNCH(k_inbound_flags, ceil_div(T, 256), 256, 0, c, ok.get(), ov.get(), cum2.get(), poff.get(), T, rb,
                   pl.inbsz.get(), delta, st.iv.get());
      }
    }
    inclusive_scan(c, st.iv.get(), st.iv.get(), V);
  }
  cas_fork.reset();  // the main stream waits for the cascade's side streams
  inclusive_scan(c, (const int64_t*)st.seq.get(), (int64_t*)st.seq.get(), V);
  // best prefix
  {
    const uint32_t blocks = std::min<uint32_t>(ceil_div(V, 256), 1024);
    DBuf<Best> partial(c, blocks), out(c, 1);
    HGP_LAUNCH(k_best_partial, blocks, 256, 0, c, (const int64_t*)st.seq.get(), st.sv.get(), st.iv.get(), V,
               partial.get());
    HGP_LAUNCH(k_best_final, 1, 256, 0, c, partial.get(), blocks, out.get());
    st.best = read_scalar(c, out.get());
  }
}

}  // namespace

void refine(Ctx& c, const RefineLevel& L, int mode, uint32_t omega, uint32_t delta, uint32_t R, bool final_counts,
            uint32_t* part, uint32_t* psizes, uint32_t* pinbound) {
  bool chainup = false;
  LevelLayout layout;
  for (uint32_t rep = 0; rep < R; ++rep) {
    RoundState st;
    analyze_round(c, L, mode, omega, delta, R, rep, chainup, part, psizes, st, &layout);
    int32_t sv = 1, iv = 1;
    int64_t gain = 0;
    bool apply = false;
    if (st.best.r != NONE) {
      sv = st.best.sv;
      iv = st.best.iv;
      gain = st.best.seq;
      apply = sv <= 0 && iv <= 0 && gain >= 0;
    }
    if (c.verbose >= 2)
      fprintf(stderr, "  refine n=%u rep=%u V=%u best=%d sv=%d iv=%d gain=%lld %s\n", L.n, rep, st.V, (int)st.best.r, sv,
              iv, (long long)gain, apply ? "apply" : "skip");
    if (apply) {
      HGP_LAUNCH(k_apply, ceil_div(L.n, 256), 256, 0, c, st.moves.get(), st.ranks.get(), L.sizes, L.n, st.best.r, part,
                 psizes, L.k);
    } else {
      // schedule of sources/refinement.cu:539-542
      if (sv > 0 && !chainup) chainup = true;
      else if (rep < R / 3) rep = R / 2;
      else if (rep < 2 * R / 3) rep = 2 * R / 3;
      else rep = R;
    }
  }
  if (final_counts && pinbound) inbound_sizes(c, L.hedges, L.hoff, L.srcc, part, L.e, L.k, pinbound);
}

uint32_t refine_analyze(Ctx& c, const RefineLevel& L, int mode, uint32_t omega, uint32_t delta, uint32_t R,
                        uint32_t fm_repeat, bool chainup, const uint32_t* part, const uint32_t* psizes,
                        const AnalyzeOut& out) {
  RoundState st;
  analyze_round(c, L, mode, omega, delta, R, fm_repeat, chainup, part, psizes, st);
  const uint32_t n = L.n, V = st.V;
  if (out.moves) copy_d2h(c, out.moves, st.moves.get(), (uint64_t)n * 4);
  if (out.gains) copy_d2h(c, out.gains, st.gains.get(), (uint64_t)n * 8);
  if (out.ranks) copy_d2h(c, out.ranks, st.ranks.get(), (uint64_t)n * 4);
  if (out.seq) {
    std::fill(out.seq, out.seq + n, 0);
    copy_d2h(c, out.seq, st.seq.get(), (uint64_t)V * 8);
  }
  if (out.sviol) {
    std::fill(out.sviol, out.sviol + n, 0);
    copy_d2h(c, out.sviol, st.sv.get(), (uint64_t)V * 4);
  }
  if (out.iviol) {
    std::fill(out.iviol, out.iviol + n, 0);
    copy_d2h(c, out.iviol, st.iv.get(), (uint64_t)V * 4);
  }
  return st.best.r;
}

uint32_t jacobi_round(Ctx& c, const RefineLevel& L, uint32_t omega, uint32_t round, uint32_t* part, uint32_t* psizes) {
  const uint32_t n = L.n;
  if (n == 0) return 0;
  PartitionLists pl;
  build_partition_lists(c, L, part, pl, false, nullptr, L.k <= MASK_K);
  DBuf<uint32_t> moves(c, n);
  DBuf<int64_t> gains(c, n);
  compute_gains(c, L, pl, HGP_MODE_KWAY, omega, round, NONE, part, psizes, moves.get(), gains.get());  // strict sizes
  DBuf<uint32_t> flag(c, n);
  DBuf<uint64_t> pos(c, (uint64_t)n + 1);
  HGP_LAUNCH(k_jacobi_movers, ceil_div(n, 256), 256, 0, c, moves.get(), gains.get(), n, flag.get());
  exclusive_scan(c, flag.get(), pos.get(), n);
  const uint32_t m = (uint32_t)read_scalar(c, pos.get() + n);
  if (m == 0) return 0;
  DBuf<uint64_t> keys(c, m);
  DBuf<uint32_t> vals(c, m), dst(c, m), before(c, std::max<uint32_t>(L.k, 1)), applied(c, 1);
  DBuf<int64_t> sz(c, m);
  DBuf<uint64_t> doff(c, (uint64_t)L.k + 1);
  HGP_LAUNCH(k_jacobi_keys, ceil_div(n, 256), 256, 0, c, flag.get(), pos.get(), gains.get(), n, keys.get(), vals.get());
  radix_sort_range(c, keys.get(), vals.get(), m);  // gain desc, node id asc (stable)
  HGP_LAUNCH(k_jacobi_dst, ceil_div(m, 256), 256, 0, c, vals.get(), moves.get(), L.sizes, m, dst.get(), sz.get());
  radix_sort(c, dst.get(), vals.get(), m, std::max(1, bit_width(L.k)));  // target asc (stable)
  HGP_LAUNCH(k_jacobi_sizes, ceil_div(m, 256), 256, 0, c, vals.get(), L.sizes, m, sz.get());
  inclusive_scan(c, sz.get(), sz.get(), m);
  segment_offsets(c, dst.get(), 0, m, L.k, doff.get());
  copy_d2d(c, before.get(), psizes, (uint64_t)L.k * 4);
  memset_async(c, applied.get(), 0, 4);
  HGP_LAUNCH(k_jacobi_apply, ceil_div(m, 256), 256, 0, c, dst.get(), vals.get(), sz.get(), doff.get(), m, L.sizes,
             before.get(), omega, part, psizes, applied.get());
  return read_scalar(c, applied.get());
}

SmallCsr small_csr(Ctx& c, const uint64_t* keys, uint32_t m, const uint64_t* hoff, const uint32_t* hedges,
                   const float* w) {
  SmallCsr r;
  r.m = m;
  if (!m) return r;
  r.rec = DBuf<uint64_t>(c, (uint64_t)m + 1);
  HGP_LAUNCH(k_small_lens, ceil_div(m, 256), 256, 0, c, keys, m, hoff, r.rec.get());
  exclusive_scan(c, r.rec.get(), r.rec.get(), m);
  const uint64_t sp = read_scalar(c, r.rec.get() + m);
  r.pins = DBuf<uint32_t>(c, sp);
  r.w = DBuf<float>(c, m);
  HGP_LAUNCH(k_small_fill, ceil_div(m, 256), 256, 0, c, keys, m, hoff, hedges, w, r.rec.get(), r.pins.get(),
             r.w.get());
  return r;
}

float time_cascade_small(Ctx& c, const SmallCsr& cs, const uint2* pm2, uint32_t block, unsigned long long* seq) {
  const uint32_t m = cs.m;
  CascadeArgs a{nullptr, nullptr, nullptr, nullptr, pm2, 0, seq};
  const SmallHedges sh{cs.rec.get(), cs.pins.get(), cs.w.get(), m};
  hipEvent_t t0, t1;
  HGP_HIP(hipEventCreate(&t0));
  HGP_HIP(hipEventCreate(&t1));
  HGP_HIP(hipEventRecord(t0, c.stream));
  HGP_LAUNCH(k_cascade_small<true>, ceil_div(m, block), block, 0, c, a, sh);
  HGP_HIP(hipEventRecord(t1, c.stream));
  HGP_HIP(hipEventSynchronize(t1));
  float ms = 0;
  HGP_HIP(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return ms;
}

void lp_refine(Ctx& c, const RefineLevel& L, uint32_t omega, uint32_t rounds, uint32_t* part, uint32_t* psizes) {
  if (rounds == 0 || L.n == 0) return;
  DBuf<uint32_t> sp(c, L.n), ss(c, std::max<uint32_t>(L.k, 1));
  int64_t cur = connectivity_fixed(c, L.hedges, L.hoff, L.w, part, L.e);
  for (uint32_t r = 0; r < rounds; ++r) {
    copy_d2d(c, sp.get(), part, (uint64_t)L.n * 4);
    copy_d2d(c, ss.get(), psizes, (uint64_t)L.k * 4);
    if (jacobi_round(c, L, omega, r, part, psizes) == 0) break;
    const int64_t now = connectivity_fixed(c, L.hedges, L.hoff, L.w, part, L.e);
    if (c.verbose >= 2) fprintf(stderr, "  lp n=%u round=%u km1 %lld -> %lld\n", L.n, r, (long long)cur, (long long)now);
    if (now > cur) {  // undo the round
      copy_d2d(c, part, sp.get(), (uint64_t)L.n * 4);
      copy_d2d(c, psizes, ss.get(), (uint64_t)L.k * 4);
      break;
    }
    cur = now;
  }
}

void inbound_sizes(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const uint32_t* srcc, const uint32_t* part,
                   uint32_t e, uint32_t k, uint32_t* out) {
  RefineLevel L{};
  L.e = e;
  L.k = k;
  L.hedges = hedges;
  L.hoff = hoff;
  L.srcc = srcc;
  PartitionLists pl;
  build_partition_lists(c, L, part, pl, true);
  copy_d2d(c, out, pl.inbsz.get(), (uint64_t)k * 4);
}

int64_t connectivity_fixed(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, const uint32_t* part,
                           uint32_t e) {
  if (e == 0) return 0;
  // partitions in use: max id + 1 (needed to size the per-partition inbound counters)
  const uint64_t P = read_scalar(c, hoff + e);
  DBuf<uint32_t> mx(c, 1);
  memset_async(c, mx.get(), 0, 4);
  // max over the partition ids of the pins
  DBuf<uint32_t> pinpart(c, P);
  HGP_LAUNCH(k_gather_u32, ceil_div(P, 256), 256, 0, c, part, hedges, (uint32_t)P, pinpart.get());
  HGP_LAUNCH(k_max_u32, std::min<uint32_t>(ceil_div(P, 256), 2048), 256, 0, c, pinpart.get(), (uint32_t)P, mx.get());
  const uint32_t k = read_scalar(c, mx.get()) + 1;
  DBuf<uint32_t> srcc0(c, e);
  memset_async(c, srcc0.get(), 0, (uint64_t)e * 4);
  RefineLevel L{};
  L.e = e;
  L.k = k;
  L.hedges = hedges;
  L.hoff = hoff;
  L.srcc = srcc0.get();
  PartitionLists pl;
  build_partition_lists(c, L, part, pl, false);
  DBuf<unsigned long long> acc(c, 1);
  memset_async(c, acc.get(), 0, 8);
  HGP_LAUNCH(k_connectivity, ceil_div(e, 256), 256, 0, c, pl.lambda.get(), w, e, acc.get());
  return (int64_t)read_scalar(c, acc.get());
}

}  // namespace hgp
