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
#include <cstdio>

#include "prim.hpp"
#include "stages.hpp"
#include "wave_iter.hpp"

namespace hgp {

namespace {

constexpr uint32_t CAND_CAP = 1024;    // LDS slots per wave
constexpr uint32_t CAND_LOG = 10;
constexpr uint32_t CAND_LIGHT = 512;   // pin-visit bound served from LDS
constexpr uint32_t CAND_LIGHT_S = 64;   // ... by the small-table instance (128 slots, 8 waves per workgroup)
constexpr uint32_t CAND_LIGHT_M = 256;  // ... by the mid-table instance (512 slots, 8 waves per workgroup)
constexpr uint32_t CAND_MED_CAP = 4096;   // LDS table slots of a medium node (48 KB)
constexpr uint32_t CAND_MED_S = 2048;     // ... of the small-table medium instance
constexpr uint32_t CAND_BIG_SLOTS = 16384;  // LDS hash slots of a big node (128 KB; half with inbound counts)
constexpr uint32_t CAND_BIG8_SLOTS = 8192;  // LDS hash slots of the 8192-slot tier (64 KB)
constexpr uint32_t CAND_DIRECT_N = 39168;   // levels up to this size index LDS by node id (<= 160 KB;
                                            // half that with inbound counts)
constexpr uint64_t HEAVY_TABLE_SLOTS = 64ull << 20;  // HBM table slots shared by the heavy grid

// Pins visited when scanning a node's touching hedges (itself excluded once per hedge).
// A node touching at most VISIT_LIGHT hedges is bounded by its own thread; the others (the
// coarse levels' nodes touch thousands) are listed for k_visit_bound_heavy.
constexpr uint32_t VISIT_LIGHT = 32, VISIT_HEAVY_GRID = 4096;
__global__ void __launch_bounds__(256) k_visit_bound(const uint64_t* __restrict__ hoff, const uint32_t* __restrict__ tch,
                                                     const uint64_t* __restrict__ toff, uint32_t n,
                                                     uint64_t* __restrict__ bound, uint32_t* __restrict__ heavy,
                                                     uint32_t* __restrict__ nheavy) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  bool to_list = false;
  if (v < n) {
    const uint64_t t0 = toff[v], t1 = toff[v + 1];
    if (t1 - t0 <= VISIT_LIGHT) {
      uint64_t b = 0;
      for (uint64_t j = t0; j < t1; ++j) {
        const uint32_t e = tch[j];
        b += hoff[e + 1] - hoff[e] - 1;
      }
      bound[v] = b;
    } else {
      to_list = true;
    }
  }
  const uint64_t m = __ballot(to_list);
  if (!m) return;
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(nheavy, (uint32_t)__popcll(m));
  base = __shfl(base, 0, 64);
  if (to_list) heavy[base + lanes_below(m)] = v;
}
// One 256-thread workgroup per listed node (persistent over the list): each thread takes every
// 256th touching hedge, four loads in flight at a time.
__global__ void __launch_bounds__(256) k_visit_bound_heavy(const uint64_t* __restrict__ hoff,
                                                           const uint32_t* __restrict__ tch,
                                                           const uint64_t* __restrict__ toff,
                                                           const uint32_t* __restrict__ heavy,
                                                           const uint32_t* __restrict__ nheavy,
                                                           uint64_t* __restrict__ bound) {
  __shared__ uint64_t s_part[4];
  const uint32_t cnt = *nheavy, wv = threadIdx.x / 64;
  for (uint32_t i = blockIdx.x; i < cnt; i += gridDim.x) {  // uniform over the workgroup
    const uint32_t v = heavy[i];
    const uint64_t t0 = toff[v], t1 = toff[v + 1];
    uint64_t b = 0;
    for (uint64_t j = t0 + threadIdx.x; j < t1; j += 4 * 256) {
      uint32_t e[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) e[u] = j + u * 256 < t1 ? tch[j + u * 256] : NONE;
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u)
        if (e[u] != NONE) b += hoff[e[u] + 1] - hoff[e[u]] - 1;
    }
    b = wave_sum(b);
    if (lane_id() == 0) s_part[wv] = b;
    __syncthreads();
    if (threadIdx.x == 0) bound[v] = s_part[0] + s_part[1] + s_part[2] + s_part[3];
    __syncthreads();
  }
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


// k >> b, 0 for b >= 64
__device__ __forceinline__ uint64_t shr64(uint64_t k, uint32_t b) { return b >= 64 ? 0ull : k >> b; }

// One wave per node of visit bound b <= CAND_LIGHT (CAP = CAND_CAP, 4 waves per workgroup),
// b <= CAND_LIGHT_M (CAP = 512, 8 waves) or b <= CAND_LIGHT_S (CAP = 128, 8 waves): smaller
// tables leave room for more waves per CU.  The
// node's LDS table has the smallest power of two >= 2b slots (at least 64), so the distinct
// neighbours (<= b) never fill it and a node with few neighbours clears and ranks few slots.
// INTER keeps the shared-inbound counts (their array is not allocated otherwise).
template <bool INTER, uint32_t CAP = CAND_CAP, uint32_t NW = 4>
__global__ void __launch_bounds__(64 * NW) k_cand_light(const uint32_t* __restrict__ hedges, const uint64_t* __restrict__ hoff,
                                                    const uint32_t* __restrict__ srcc, const uint32_t* __restrict__ tch,
                                                    const uint64_t* __restrict__ toff, const uint32_t* __restrict__ inb,
                                                    const float* __restrict__ w, const uint32_t* __restrict__ sizes,
                                                    const uint32_t* __restrict__ list, uint32_t nlist,
                                                    const uint64_t* __restrict__ bound, uint32_t cnc, uint32_t omega,
                                                    uint32_t delta, uint32_t* __restrict__ pairs,
                                                    uint32_t* __restrict__ scores) {
  constexpr uint32_t SLOTS = CAP / 64;
  __shared__ uint32_t s_keys[NW][CAP];
  __shared__ uint32_t s_score[NW][CAP];
  __shared__ uint32_t s_inter[INTER ? NW : 1][INTER ? CAP : 1];
  const uint32_t wv = threadIdx.x / 64, l = lane_id();
  const uint32_t idx = xcd_block(blockIdx.x, gridDim.x) * NW + wv;
  if (idx >= nlist) return;
  const uint32_t u = list[idx];
  uint32_t* keys = s_keys[wv];
  uint32_t* score = s_score[wv];
  uint32_t* inter = INTER ? s_inter[wv] : nullptr;
  uint32_t cap = 64;
  while (cap < 2 * bound[u] && cap < CAP) cap <<= 1;
  for (uint32_t s = l; s < cap; s += 64) {
    keys[s] = NONE;
    score[s] = 0;
    if (INTER) inter[s] = 0;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const uint64_t t0 = toff[u], t1 = toff[u + 1];
  const uint32_t inb_u = inb[u];
  // per touching hedge, once: its contribution floor(W / |e|) and its source-pin count
  auto hvf = [&](uint32_t e, uint32_t len) { return make_uint2(fixed_weight(w[e]) / len, INTER ? srcc[e] : 0u); };
  wave_visit_pins_hv_until(tch, t0, t1, hoff, hedges, hvf, [&](const PinVisitHV& pv) {
    if (!pv.act || pv.pin == u) return;
    const uint32_t contrib = pv.hv.x;
    const bool shared_in = INTER && (pv.j - t0) < inb_u && pv.pos >= pv.hv.y;
    uint32_t h = slot_hash(pv.pin) & (cap - 1);
    for (;;) {
      const uint32_t old = atomicCAS(&keys[h], NONE, pv.pin);
      if (old == NONE || old == pv.pin) break;
      h = (h + 1) & (cap - 1);
    }
    atomicAdd(&score[h], contrib);
    if (INTER && shared_in) atomicAdd(&inter[h], 1u);
  }, [] { return false; });
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const uint32_t size_u = sizes[u];
  uint64_t cand[SLOTS];
#pragma unroll
  for (uint32_t s = 0; s < SLOTS; ++s) {
    const uint32_t slot = l + 64 * s;
    const uint32_t v = slot < cap ? keys[slot] : NONE;
    cand[s] = v == NONE ? 0ull
                        : cand_key(u, v, score[slot], INTER ? inter[slot] : 0u, size_u, inb_u, sizes, inb, omega, delta);
  }
  // The top cnc keys by a radix select (as block_select_write, for one wave): the keys are in
  // registers now, so the table's score array holds the histogram (CAP >= 128 bins) and its
  // key array the selected keys.
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  constexpr uint32_t DB = CAP >= 256 ? 8 : 7, BINS = 1u << DB, BPL = BINS / 64;  // digit bits, bins per lane
  static_assert(CAP >= BINS, "the histogram fits the score array");
  uint32_t* hist = score;
  uint64_t m = 0;
#pragma unroll
  for (uint32_t s = 0; s < SLOTS; ++s) m = cand[s] > m ? cand[s] : m;
  const uint64_t kmax = wave_max(m);
  uint64_t thr = ~0ull;  // the selected keys are the non-zero ones >= thr
  if (kmax != 0) {
    uint32_t upper = 64 - (uint32_t)__clzll((long long)kmax), need = cnc;
    uint64_t prefix = 0;
    for (;;) {
      const uint32_t lower = upper > DB ? upper - DB : 0u;
      const uint32_t dmask = (1u << (upper - lower)) - 1u;
#pragma unroll
      for (uint32_t i = 0; i < BPL; ++i) hist[l + 64 * i] = 0;
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
      for (uint32_t s = 0; s < SLOTS; ++s) {
        const uint64_t k = cand[s];
        if (k && shr64(k, upper) == prefix) atomicAdd(&hist[(uint32_t)(k >> lower) & dmask], 1u);
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      uint32_t c[BPL], cl = 0;  // lane l: bins BPL * l ..
#pragma unroll
      for (uint32_t i = 0; i < BPL; ++i) {
        c[i] = hist[BPL * l + i];
        cl += c[i];
      }
      const uint32_t incl = wave_scan_incl(cl);
      const uint32_t total = __shfl(incl, 63, 64);
      if (total <= need) {  // only on the first pass: every candidate is taken
        thr = 0;
        break;
      }
      const uint32_t at_or_above = total - (incl - cl);
      const uint64_t ok = __ballot(at_or_above >= need);  // lanes 0 .. L
      const uint32_t L = 63u - (uint32_t)__clzll((long long)ok);
      uint32_t cum = at_or_above - cl, b = 0, cb = c[0];  // lane L's bins, from the top
#pragma unroll
      for (int i = (int)BPL - 1; i >= 0; --i) {
        if (cum + c[i] >= need) {
          b = (uint32_t)i;
          cb = c[i];
          break;
        }
        cum += c[i];
      }
      const uint32_t bL = (uint32_t)__builtin_amdgcn_readlane((int)b, (int)L);
      const uint32_t cumL = (uint32_t)__builtin_amdgcn_readlane((int)cum, (int)L);
      const uint32_t cbL = (uint32_t)__builtin_amdgcn_readlane((int)cb, (int)L);
      prefix = (prefix << (upper - lower)) | (BPL * L + bL);
      need -= cumL;
      if (cbL == need || lower == 0) {  // every key of the bin is taken (a bin is one key at bit 0)
        thr = prefix << lower;
        break;
      }
      upper = lower;
      __builtin_amdgcn_wave_barrier();  // the bins are read before the next pass clears them
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }
  // the selected keys (at most cnc) into the key array as (lo, hi) pairs, then ranked
  uint32_t mine = 0;
#pragma unroll
  for (uint32_t s = 0; s < SLOTS; ++s) mine += (cand[s] && cand[s] >= thr) ? 1u : 0u;
  const uint32_t pincl = wave_scan_incl(mine);
  const uint32_t n = min(__shfl(pincl, 63, 64), cnc);
  uint32_t at = pincl - mine;
#pragma unroll
  for (uint32_t s = 0; s < SLOTS; ++s) {
    const uint64_t k = cand[s];
    if (k && k >= thr) {
      if (at < 16) {
        keys[2 * at] = (uint32_t)k;
        keys[2 * at + 1] = (uint32_t)(k >> 32);
      }
      ++at;
    }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (l < cnc) {
    if (l < n) {
      const uint64_t k = ((uint64_t)keys[2 * l + 1] << 32) | keys[2 * l];
      uint32_t r = 0;
      for (uint32_t j = 0; j < n; ++j) r += (((uint64_t)keys[2 * j + 1] << 32) | keys[2 * j]) > k ? 1u : 0u;
      pairs[(uint64_t)u * cnc + r] = (uint32_t)k;
      scores[(uint64_t)u * cnc + r] = (uint32_t)(k >> 32);
    } else {
      pairs[(uint64_t)u * cnc + l] = NONE;
      scores[(uint64_t)u * cnc + l] = 0;
    }
  }
}

// Sorted (descending) per-thread list of the TOPN best candidate keys seen so far.  A thread
// that owns at most TOPN < 16 table slots keeps all of them in a list of that length (fewer
// registers during the ranking, so more waves per SIMD during the traversal).
template <int TOPN>
__device__ __forceinline__ void top_insert(uint64_t (&top)[TOPN], uint64_t k) {
  if (k <= top[TOPN - 1]) return;
#pragma unroll
  for (int i = 0; i < TOPN; ++i) {
    const bool sw = k > top[i];
    const uint64_t t = top[i];
    top[i] = sw ? k : t;
    k = sw ? t : k;
  }
}

// Block-wide top-cnc of the per-thread lists (keys are distinct: they embed the node id).
// Each of the NW waves extracts its cnc best by repeated wave maxima; wave 0 then merges the
// NW x 16 survivors (NW / 4 per lane).
template <int NW, int TOPN>
__device__ __forceinline__ void block_top_write(uint64_t (&top)[TOPN], uint32_t cnc, uint64_t* s_w, uint32_t u,
                                                uint32_t* __restrict__ pairs, uint32_t* __restrict__ scores) {
  static_assert(NW % 4 == 0, "NW must be a multiple of 4");
  const uint32_t wv = threadIdx.x / 64, l = lane_id();
  for (uint32_t i = 0; i < 16; ++i) {
    uint64_t m = 0;
    if (i < cnc) {
      m = wave_max(top[0]);
      if (m != 0 && top[0] == m) {
#pragma unroll
        for (int j = 0; j < TOPN - 1; ++j) top[j] = top[j + 1];
        top[TOPN - 1] = 0;
      }
    }
    if (l == 0) s_w[wv * 16 + i] = m;
  }
  __syncthreads();
  if (wv == 0) {
    constexpr int PER = NW / 4;
    uint64_t x[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) x[q] = s_w[l + 64 * q];
    for (uint32_t i = 0; i < cnc; ++i) {
      uint64_t lm = 0;
#pragma unroll
      for (int q = 0; q < PER; ++q) lm = x[q] > lm ? x[q] : lm;
      const uint64_t m = wave_max(lm);
#pragma unroll
      for (int q = 0; q < PER; ++q)
        if (m != 0 && x[q] == m) x[q] = 0;
      if (l == 0) {
        pairs[(uint64_t)u * cnc + i] = m ? (uint32_t)m : NONE;
        scores[(uint64_t)u * cnc + i] = (uint32_t)(m >> 32);
      }
    }
  }
  __syncthreads();
}

// Exact top-cnc (cnc <= 16) of the candidate keys of a node's LDS table, written to the
// node's pairs/scores in descending key order (NONE / 0 past the candidates).  Slot s holds
// its key's high word in hi[s] (0: no candidate) and its low word (the neighbour id) in
// keys[s], or s itself when the table is indexed by node id.  Keys are distinct (they embed
// the neighbour id).  A radix select: each pass histograms the next 8 bits below those fixed
// so far, over the keys that share the fixed bits, and fixes the bin that holds the cnc-th
// largest key; it stops once every key of that bin is taken.  The first pass starts below
// the largest key's top bit.  The passes sweep the table in parallel, where selecting by
// repeated wave maxima would be a chain of dependent reductions.
struct SelectScratch {
  uint32_t hist[256];
  uint64_t sel[16];    // the selected keys; first the per-wave maxima
  uint64_t prefix;     // bits of the cnc-th key fixed so far (above bit `upper`)
  uint32_t need;       // keys still to take among those sharing the prefix
  uint32_t done;
  uint32_t nsel;
  uint32_t pad;
};
constexpr uint32_t SELECT_WORDS = (sizeof(SelectScratch) + 7) / 8;  // 8-byte words of the scratch
static_assert(SELECT_WORDS <= 256, "the 1024-thread kernels' 256-word scratch holds a SelectScratch");
template <int NW>
__device__ __forceinline__ void block_select_write(uint32_t cap, bool direct, const uint32_t* keys, const uint32_t* hi,
                                                   uint64_t kmax, uint32_t cnc, uint32_t u,
                                                   uint32_t* __restrict__ pairs, uint32_t* __restrict__ scores,
                                                   SelectScratch* ss) {
  static_assert(NW <= 16 && 64 * NW >= 256, "one histogram bin per thread at most, at most 16 wave maxima");
  constexpr uint32_t B = 64 * NW;
  const uint32_t tid = threadIdx.x, l = lane_id(), wv = tid / 64;
  auto key_at = [&](uint32_t s) -> uint64_t {
    const uint32_t h = hi[s];
    return h ? ((uint64_t)h << 32) | (direct ? s : keys[s]) : 0ull;
  };
  kmax = wave_max(kmax);
  if (l == 0) ss->sel[wv] = kmax;
  if (tid == 0) {
    ss->prefix = 0;
    ss->need = cnc;
    ss->done = 0;
    ss->nsel = 0;
  }
  __syncthreads();
  uint64_t bmax = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) bmax = ss->sel[w] > bmax ? ss->sel[w] : bmax;
  uint64_t thr = ~0ull;  // the selected keys are the non-zero ones >= thr
  if (bmax != 0) {       // uniform
    uint32_t upper = 64 - (uint32_t)__clzll((long long)bmax);  // every key is below 2^upper
    for (;;) {
      const uint32_t lower = upper > 8 ? upper - 8 : 0u;
      __syncthreads();  // the previous pass's reads of the histogram and state are done
      if (tid < 256) ss->hist[tid] = 0;
      __syncthreads();
      const uint64_t pre = ss->prefix;
      const uint32_t dmask = (1u << (upper - lower)) - 1u;
      for (uint32_t s = tid; s < cap; s += B) {
        const uint64_t k = key_at(s);
        if (k && shr64(k, upper) == pre) atomicAdd(&ss->hist[(uint32_t)(k >> lower) & dmask], 1u);
      }
      __syncthreads();
      if (wv == 0) {  // lane l holds bins 4l .. 4l + 3
        const uint32_t c0 = ss->hist[4 * l], c1 = ss->hist[4 * l + 1], c2 = ss->hist[4 * l + 2],
                       c3 = ss->hist[4 * l + 3];
        const uint32_t c4 = c0 + c1 + c2 + c3;
        const uint32_t incl = wave_scan_incl(c4);
        const uint32_t total = __shfl(incl, 63, 64);
        const uint32_t need = ss->need;
        if (total <= need) {  // only on the first pass: every candidate is taken
          if (l == 0) {
            ss->prefix = 0;
            ss->done = 1;
          }
        } else {
          const uint32_t at_or_above = total - (incl - c4);  // keys in bins >= 4l
          const uint64_t ok = __ballot(at_or_above >= need);  // lanes 0 .. L (lane 0: total > need)
          const uint32_t L = 63u - (uint32_t)__clzll((long long)ok);
          if (l == L) {
            uint32_t cum = at_or_above - c4, b, cb;  // cum: keys in the bins above bin 4L + b
            if (cum + c3 >= need) b = 3, cb = c3;
            else if (cum += c3, cum + c2 >= need) b = 2, cb = c2;
            else if (cum += c2, cum + c1 >= need) b = 1, cb = c1;
            else cum += c1, b = 0, cb = c0;
            const uint32_t nn = need - cum;
            ss->prefix = (pre << (upper - lower)) | (4 * L + b);
            ss->need = nn;
            ss->done = cb == nn ? 1u : 0u;
          }
        }
      }
      __syncthreads();
      if (ss->done || lower == 0) {  // uniform; with lower == 0 a bin is one key: done
        thr = ss->prefix << lower;
        break;
      }
      upper = lower;
    }
  }
  for (uint32_t s = tid; s < cap; s += B) {
    const uint64_t k = key_at(s);
    if (k && k >= thr) {
      const uint32_t i = atomicAdd(&ss->nsel, 1u);
      if (i < 16) ss->sel[i] = k;
    }
  }
  __syncthreads();
  const uint32_t n = ss->nsel < 16 ? ss->nsel : 16u;
  if (tid < cnc) {
    if (tid < n) {  // rank among the selected keys
      const uint64_t k = ss->sel[tid];
      uint32_t r = 0;
      for (uint32_t j = 0; j < n; ++j) r += ss->sel[j] > k ? 1u : 0u;
      pairs[(uint64_t)u * cnc + r] = (uint32_t)k;
      scores[(uint64_t)u * cnc + r] = (uint32_t)(k >> 32);
    } else {
      pairs[(uint64_t)u * cnc + tid] = NONE;
      scores[(uint64_t)u * cnc + tid] = 0;
    }
  }
  __syncthreads();  // the table and the scratch are reused by the next node
}

// Table capacity of a node whose touching hedges hold `bound` other pins: distinct
// neighbours are at most min(bound, n - 1); a power-of-two open-addressing table of twice
// that (plus slack) never fills.
__device__ __forceinline__ uint32_t hash_cap(uint64_t bound, uint32_t n) {
  const uint64_t d = bound < (uint64_t)n ? bound : (uint64_t)n;
  uint64_t c = 64;
  while (c < 2 * d + 64) c <<= 1;
  return (uint32_t)c;
}

// Node classes: 0 light (one wave, LDS 1024 slots), 1 medium (256 threads, LDS table of
// <= 4096 slots), 2 big (1024 threads, one LDS table: indexed by node id when the level has
// at most direct_n nodes, else hashed with <= big_cap slots), 3 heavy (HBM tables).
// cap[v] = table slots; a heavy node whose table would not be smaller than n indexes directly.
__global__ void k_cand_tiers(const uint64_t* __restrict__ bound, uint32_t n, uint32_t direct_n, uint32_t big8_cap,
                             uint32_t big_cap, uint32_t* __restrict__ f_small, uint32_t* __restrict__ f_mid,
                             uint32_t* __restrict__ f_light, uint32_t* __restrict__ f_med_s,
                             uint32_t* __restrict__ f_med, uint32_t* __restrict__ f_big8, uint32_t* __restrict__ f_big,
                             uint32_t* __restrict__ f_direct, uint32_t* __restrict__ cap) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint64_t b = bound[v];
  const uint32_t c = hash_cap(b, n);
  const bool light = b <= CAND_LIGHT;
  const bool med = !light && c <= CAND_MED_CAP;
  const bool big8 = !light && !med && c <= big8_cap;
  const bool big = !light && !med && !big8 && c <= big_cap;                // hashed, on any level
  const bool direct = !light && !med && !big8 && !big && n <= direct_n;  // indexed by node id
  f_small[v] = b <= CAND_LIGHT_S ? 1u : 0u;
  f_mid[v] = b > CAND_LIGHT_S && b <= CAND_LIGHT_M ? 1u : 0u;
  f_light[v] = light && b > CAND_LIGHT_M ? 1u : 0u;
  f_med_s[v] = med && c <= CAND_MED_S ? 1u : 0u;
  f_med[v] = med && c > CAND_MED_S ? 1u : 0u;
  f_big8[v] = big8 ? 1u : 0u;
  f_big[v] = big ? 1u : 0u;
  f_direct[v] = direct ? 1u : 0u;
  // indexed by node id: the direct tier, and heavy nodes whose table would not be smaller than n;
  // every other tier hashes into a power-of-two table of c slots
  const bool heavy = !light && !med && !big8 && !big && !direct;
  cap[v] = direct || (heavy && (uint64_t)n <= c) ? n : c;
}

// node lists of the tiers flagged in f[0..T-2] (exclusive scans p[0..T-2]); the rest go to
// the last list
constexpr int TIERS = 9;
struct TierLists {
  const uint32_t* f[TIERS - 1];
  const uint64_t* p[TIERS - 1];
  uint32_t* l[TIERS];
};
__global__ void k_split_tiers(TierLists t, uint32_t n) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  uint64_t before = 0;
#pragma unroll
  for (int i = 0; i < TIERS - 1; ++i) {
    if (t.f[i][v]) {
      t.l[i][t.p[i][v]] = v;
      return;
    }
    before += t.p[i][v];
  }
  t.l[TIERS - 1][v - before] = v;
}

// Largest table among the listed nodes, each capped at `try_cap`.
__global__ void k_heavy_max_cap(const uint32_t* __restrict__ heavy, const uint32_t* __restrict__ cap, uint32_t nh,
                                uint32_t try_cap, uint32_t* out) {
  uint32_t m = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nh; i += gridDim.x * blockDim.x) {
    const uint32_t c = cap[heavy[i]];
    const uint32_t r = c < try_cap ? c : try_cap;
    m = r > m ? r : m;
  }
  m = wave_max(m);
  if (lane_id() == 0) atomicMax(out, m);
}

// Scores one node with a whole workgroup of NW waves: the waves split its touching hedges
// and insert into the table, then every entry is ranked.  The table is LDS or the block's
// HBM region (GMEM).  Hashed: keys[] holds node ids.  Direct (cap == level size): slot = node
// id and keys[] is a presence bitmap.  INTER keeps the shared-inbound counts that the
// inbound-union filter needs; without it (Delta >= twice the hedge count) they are 0.
//
// OPT (optimistic): the node's distinct neighbours may exceed the table.  New keys are
// counted; past `limit` the workgroup stops and the caller hands the node to a larger table.  At most one insert per thread is in flight when
// the limit trips, so the table never fills.  Returns false when the node overflowed.
// TOPN: length of the per-thread candidate lists, 16 unless every thread owns at most TOPN
// table slots (then its list holds all of them).
template <bool GMEM, bool INTER, int NW, bool OPT = false, bool P2 = true, int CH = 1, bool PIPE = true,
          int TOPN = 16>
__device__ __forceinline__ bool cand_block_node(uint32_t u, uint32_t cap, bool direct, uint32_t* keys, uint32_t* score,
                                                uint32_t* inter, const uint32_t* __restrict__ hedges,
                                                const uint64_t* __restrict__ hoff, const uint32_t* __restrict__ srcc,
                                                const uint32_t* __restrict__ tch, const uint64_t* __restrict__ toff,
                                                const uint32_t* __restrict__ inb, const float* __restrict__ w,
                                                const uint32_t* __restrict__ sizes, uint32_t cnc, uint32_t omega,
                                                uint32_t delta, uint32_t* __restrict__ pairs,
                                                uint32_t* __restrict__ scores, uint64_t* s_w,
                                                uint32_t* s_state = nullptr, uint32_t limit = 0) {
  const uint32_t nkeys = direct ? (cap + 31) / 32 : cap;
  const uint64_t t0 = toff[u], t1 = toff[u + 1];
  const uint32_t inb_u = inb[u];
  const uint32_t size_u = sizes[u];
  const uint32_t wv = threadIdx.x / 64;
  const uint64_t cnt = t1 - t0;
  const uint64_t a = t0 + cnt * wv / NW, b = t0 + cnt * (wv + 1) / NW;
  auto overflowed = [&] {
    return OPT && __hip_atomic_load(&s_state[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u;
  };
  // LDS tables scored in one pass are ranked by a radix select (s_w is a SelectScratch);
  // tables in HBM, or scored in several passes, by per-thread candidate lists
  constexpr bool SELECT = !GMEM && CH == 1;
  uint64_t top[TOPN];
#pragma unroll
  for (int i = 0; i < TOPN; ++i) top[i] = 0;
  // CH > 1: the neighbours are split by id modulo CH and scored one class per pass, so the
  // table holds a CH-th of them; the per-thread top lists carry over between passes.
  for (int chunk = 0; chunk < CH; ++chunk) {
    for (uint32_t s = threadIdx.x; s < cap; s += blockDim.x) {
      score[s] = 0;
      if (INTER) inter[s] = 0;
    }
    for (uint32_t s = threadIdx.x; s < nkeys; s += blockDim.x) keys[s] = direct ? 0u : NONE;
    if (OPT && threadIdx.x == 0) {
      s_state[0] = 0;  // new keys
      s_state[1] = 0;  // overflow flag
    }
    if (GMEM) __threadfence();
    __syncthreads();
    auto hvf = [&](uint32_t e, uint32_t len) { return make_uint2(fixed_weight(w[e]) / len, INTER ? srcc[e] : 0u); };
    wave_visit_pins_hv_until<PIPE>(tch, a, b, hoff, hedges, hvf, [&](const PinVisitHV& pv) {
      if (!pv.act || pv.pin == u) return;
      if (CH > 1 && pv.pin % CH != (uint32_t)chunk) return;
      const uint32_t contrib = pv.hv.x;  // floor(W / |e|), computed once per touching hedge
      const bool shared_in = INTER && (pv.j - t0) < inb_u && pv.pos >= pv.hv.y;
      if (direct) {
        const uint32_t bit = 1u << (pv.pin & 31u);
        if (!(atomic_load_relaxed(&keys[pv.pin >> 5]) & bit)) atomicOr(&keys[pv.pin >> 5], bit);
        atomicAdd(&score[pv.pin], contrib);
        if (INTER && shared_in) atomicAdd(&inter[pv.pin], 1u);
      } else {
        // P2: power-of-two table, masked index; otherwise any size, multiply-high index
        uint32_t h = P2 ? slot_hash(pv.pin) & (cap - 1) : __umulhi(slot_hash(pv.pin), cap);
        for (;;) {
          const uint32_t old = atomicCAS(&keys[h], NONE, pv.pin);
          if (old == NONE) {
            if (OPT && atomicAdd(&s_state[0], 1u) + 1u > limit) atomicOr(&s_state[1], 1u);
            break;
          }
          if (old == pv.pin) break;
          h = P2 ? (h + 1) & (cap - 1) : (h + 1 == cap ? 0u : h + 1);
        }
        atomicAdd(&score[h], contrib);
        if (INTER && shared_in) atomicAdd(&inter[h], 1u);
      }
    }, overflowed);
    if (GMEM) __threadfence();
    __syncthreads();
    if (OPT && s_state[1]) return false;
    if constexpr (SELECT) {
      // each slot's key, its high word written over the score (0: no candidate)
      uint64_t kmax = 0;
      for (uint32_t s = threadIdx.x; s < cap; s += blockDim.x) {
        uint32_t v;
        if (direct) {
          if (!((keys[s >> 5] >> (s & 31u)) & 1u)) continue;
          v = s;
        } else {
          v = keys[s];
          if (v == NONE) continue;
        }
        const uint64_t k = cand_key(u, v, score[s], INTER ? inter[s] : 0u, size_u, inb_u, sizes, inb, omega, delta);
        score[s] = (uint32_t)(k >> 32);
        kmax = k > kmax ? k : kmax;
      }
      block_select_write<NW>(cap, direct, keys, score, kmax, cnc, u, pairs, scores, (SelectScratch*)s_w);
      return true;
    }
    for (uint32_t s = threadIdx.x; s < cap; s += blockDim.x) {
      uint32_t v;
      if (direct) {
        if (!((keys[s >> 5] >> (s & 31u)) & 1u)) continue;
        v = s;
      } else {
        v = keys[s];
        if (v == NONE) continue;
      }
      top_insert(top, cand_key(u, v, score[s], INTER ? inter[s] : 0u, size_u, inb_u, sizes, inb, omega, delta));
    }
    // the pin ids visited; per touching hedge its list entry, offsets, weight (and source count)
    if (CH > 1) __syncthreads();  // the next pass clears the table
  }
  block_top_write<NW, TOPN>(top, cnc, s_w, u, pairs, scores);
  return true;
}

#define CAND_ARGS                                                                                                   \
  const uint32_t *__restrict__ hedges, const uint64_t *__restrict__ hoff, const uint32_t *__restrict__ srcc,        \
      const uint32_t *__restrict__ tch, const uint64_t *__restrict__ toff, const uint32_t *__restrict__ inb,        \
      const float *__restrict__ w, const uint32_t *__restrict__ sizes, const uint32_t *__restrict__ list,           \
      const uint32_t *__restrict__ caps, uint32_t cnc, uint32_t omega, uint32_t delta, uint32_t *__restrict__ pairs, \
      uint32_t *__restrict__ scores

// CAP: the LDS table (CAND_MED_CAP, or CAND_MED_S for nodes whose table fits it: half the LDS,
// so twice the workgroups per CU)
// NW waves per node: 8 for the larger table, whose LDS allows only four workgroups per CU
template <bool INTER, uint32_t CAP = CAND_MED_CAP, int NW = 4>
__global__ void __launch_bounds__(64 * NW) k_cand_medium(CAND_ARGS) {
  __shared__ uint32_t s_keys[CAP];
  __shared__ uint32_t s_score[CAP];
  __shared__ uint32_t s_inter[INTER ? CAP : 1];
  __shared__ uint64_t s_w[SELECT_WORDS];
  const uint32_t u = list[xcd_block(blockIdx.x, gridDim.x)];
  cand_block_node<false, INTER, NW>(u, caps[u], false, s_keys, s_score, s_inter, hedges, hoff, srcc, tch, toff, inb, w,
                                    sizes, cnc, omega, delta, pairs, scores, s_w);
}

// One node per 1024-thread workgroup with a hashed LDS table of up to CAND_BIG_SLOTS.
template <bool INTER>
__global__ void __launch_bounds__(1024) k_cand_big(CAND_ARGS) {
  constexpr uint32_t SLOTS = INTER ? CAND_BIG_SLOTS / 2 : CAND_BIG_SLOTS;
  __shared__ uint32_t s_keys[SLOTS];
  __shared__ uint32_t s_score[SLOTS];
  __shared__ uint32_t s_inter[INTER ? SLOTS : 1];
  __shared__ uint64_t s_w[256];
  const uint32_t u = list[xcd_block(blockIdx.x, gridDim.x)];
  cand_block_node<false, INTER, 16>(u, caps[u], false, s_keys, s_score, s_inter, hedges, hoff, srcc, tch, toff, inb,
                                    w, sizes, cnc, omega, delta, pairs, scores, s_w);
}

// One node per 1024-thread workgroup with a hashed LDS table of CAND_BIG8_SLOTS (64 KB, no
// shared-inbound counts): two workgroups per CU where the big tier's table leaves room for one.
__global__ void __launch_bounds__(1024) k_cand_big8(CAND_ARGS) {
  __shared__ uint32_t s_keys[CAND_BIG8_SLOTS];
  __shared__ uint32_t s_score[CAND_BIG8_SLOTS];
  __shared__ uint32_t s_inter[1];
  __shared__ uint64_t s_w[256];
  const uint32_t u = list[xcd_block(blockIdx.x, gridDim.x)];
  cand_block_node<false, false, 16>(u, caps[u], false, s_keys, s_score, s_inter, hedges, hoff, srcc, tch, toff, inb, w,
                                    sizes, cnc, omega, delta, pairs, scores, s_w);
}

// Nodes whose worst-case table exceeds LDS: try the largest LDS table that fits (their
// distinct neighbours are usually far fewer than their visit bound); overflowing nodes are
// listed in `ovf` for the HBM-table kernel.  The table is not a power of two (it fills the
// 160 KB of LDS), and the attempt stops at 3/4 load: with at most one insert per thread in
// flight after that, LIMIT + 1024 < SLOTS keeps a free slot for every probe.
constexpr uint32_t OPT_SLOTS_INTER = 13440;  // 12 bytes per slot
constexpr uint32_t OPT_SLOTS = 20160;        // 8 bytes per slot
constexpr int OPT_CHUNKS = 4;                // passes of the second LDS attempt
// CH > 1 scores the neighbours in CH passes over the touching hedges (one id class each),
// for nodes that overflowed the single-pass table.
template <bool INTER, int CH>
__global__ void __launch_bounds__(1024) k_cand_opt(CAND_ARGS, uint32_t* __restrict__ ovf,
                                                   uint32_t* __restrict__ ovf_count) {
  constexpr uint32_t SLOTS = INTER ? OPT_SLOTS_INTER : OPT_SLOTS;
  constexpr uint32_t LIMIT = SLOTS / 4 * 3;
  static_assert(LIMIT + 1024 < SLOTS, "table must keep a free slot for every in-flight insert");
  __shared__ uint32_t s_keys[SLOTS];
  __shared__ uint32_t s_score[SLOTS];
  __shared__ uint32_t s_inter[INTER ? SLOTS : 1];
  __shared__ uint64_t s_w[256];
  __shared__ uint32_t s_state[2];
  const uint32_t u = list[xcd_block(blockIdx.x, gridDim.x)];
  const bool done = cand_block_node<false, INTER, 16, true, false, CH>(u, SLOTS, false, s_keys, s_score, s_inter,
                                                                       hedges, hoff, srcc, tch, toff, inb, w, sizes,
                                                                       cnc, omega, delta, pairs, scores, s_w, s_state,
                                                                       LIMIT);
  if (!done && threadIdx.x == 0) ovf[atomicAdd(ovf_count, 1u)] = u;
}

// One node per 1024-thread workgroup, LDS score array indexed by node id (levels of at most
// CAND_DIRECT_N nodes, half that with shared-inbound counts).
template <bool INTER>
__global__ void __launch_bounds__(1024) k_cand_direct(CAND_ARGS) {
  constexpr uint32_t N = INTER ? CAND_DIRECT_N / 2 : CAND_DIRECT_N;
  __shared__ uint32_t s_bits[N / 32];
  __shared__ uint32_t s_score[N];
  __shared__ uint32_t s_inter[INTER ? N : 1];
  __shared__ uint64_t s_w[256];
  const uint32_t u = list[xcd_block(blockIdx.x, gridDim.x)];
  // without the pipelined traversal: with it this kernel measured no faster (its occupancy is
  // set by the LDS score array)
  cand_block_node<false, INTER, 16, false, true, 1, false>(u, caps[u], true, s_bits, s_score, s_inter, hedges, hoff,
                                                           srcc, tch, toff, inb, w, sizes, cnc, omega, delta, pairs,
                                                           scores, s_w);
}

// Heavy nodes: a persistent grid, each workgroup owning `region` slots of HBM tables and
// taking nodes blockIdx.x, blockIdx.x + gridDim.x, ...  A table costs its size to clear and
// to rank, and the worst-case size (from the visit bound) can be far above the distinct
// neighbours, so nodes whose worst case exceeds `try_cap` first try a table of try_cap
// slots (optimistically, stopping at `try_fill` entries) and only the ones that overflow it go to a
// second launch (try_cap = 0) with their worst-case table.  Nodes indexed by node id
// (cap == n) also wait for the second launch, so that they do not size every region.
__global__ void __launch_bounds__(256) k_cand_heavy(CAND_ARGS, uint32_t nlist, uint32_t n, uint64_t region,
                                                    uint32_t* __restrict__ gkeys, uint32_t* __restrict__ gscore,
                                                    uint32_t* __restrict__ ginter, uint32_t try_cap, uint32_t try_fill,
                                                    uint32_t* __restrict__ ovf, uint32_t* __restrict__ ovf_count) {
  __shared__ uint64_t s_w[64];
  __shared__ uint32_t s_state[2];
  uint32_t* keys = gkeys + region * blockIdx.x;
  uint32_t* score = gscore + region * blockIdx.x;
  uint32_t* inter = ginter + region * blockIdx.x;
  for (uint32_t i = blockIdx.x; i < nlist; i += gridDim.x) {
    const uint32_t u = list[i];
    const uint32_t cap = caps[u];
    if (try_cap == 0 || cap <= try_cap) {
      cand_block_node<true, true, 4>(u, cap, cap == n, keys, score, inter, hedges, hoff, srcc, tch, toff, inb, w,
                                     sizes, cnc, omega, delta, pairs, scores, s_w);
    } else if (cap == n) {
      if (threadIdx.x == 0) ovf[atomicAdd(ovf_count, 1u)] = u;
    } else {
      const bool done = cand_block_node<true, true, 4, true>(u, try_cap, false, keys, score, inter, hedges, hoff, srcc,
                                                             tch, toff, inb, w, sizes, cnc, omega, delta, pairs,
                                                             scores, s_w, s_state, try_fill);
      if (!done && threadIdx.x == 0) ovf[atomicAdd(ovf_count, 1u)] = u;
      __syncthreads();  // s_state is reused by the next node
    }
  }
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
// S and best are only ever changed by atomic read-modify-writes, which are performed at the
// device's coherence point, and read the same way (adding 0 / or-ing 0).  A child waits for
// its two updates to complete before it decrements the parent's pending count, and the last
// child reads the parent's values after that decrement returned, so it sees every child's
// contribution without device-wide cache write-backs or invalidations.
__device__ void sweep_up(uint32_t cur, const uint32_t* __restrict__ parent, const uint32_t* __restrict__ wgt,
                         uint32_t* pending, unsigned long long* S, unsigned long long* best, uint32_t* processed) {
  for (;;) {
    const uint64_t Sv = atomicAdd(&S[cur], 0ull);
    const uint64_t bv = atomicOr(&best[cur], 0ull);
    const uint64_t bg = bv >> 32;
    const uint64_t wout = Sv + bg;
    __hip_atomic_store(&processed[cur], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t p = parent[cur];
    if (p == NONE) return;
    const int64_t gain = (int64_t)wgt[cur] - (int64_t)bg;
    const uint64_t key = ((uint64_t)(gain > 0 ? gain : 0) << 32) | cur;
    const uint64_t r1 = atomicAdd(&S[p], (unsigned long long)wout);
    const uint64_t r2 = atomicMax(&best[p], (unsigned long long)key);
    // both updates have returned (so they are performed) before the count goes down: the
    // empty asm consumes the returned values, which keeps them returning atomics
    asm volatile("" ::"v"(r1), "v"(r2));
    __builtin_amdgcn_s_waitcnt(0);
    const uint32_t old = atomicSub(&pending[p], 1u);
    if (old != 1u) return;
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
  parent[v] = NONE;  // read by no other thread of this launch
  const uint32_t old = atomicSub(&pending[p], 1u);
  if (old == 1u) sweep_up(p, parent, wgt, pending, S, best, processed);
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
  HGP_REQUIRE(cnc >= 1 && cnc <= 16, HGP_E_INVALID, "candidates: candidates_count must be in 1..16");
  if (n == 0) return;
  // shared-inbound counts matter only when two inbound counts (each <= e) can exceed Delta
  const bool inter = (uint64_t)delta < 2ull * e;
  const uint32_t direct_n = inter ? CAND_DIRECT_N / 2 : CAND_DIRECT_N;
  const uint32_t big_cap = inter ? CAND_BIG_SLOTS / 2 : CAND_BIG_SLOTS;
  const uint32_t g = ceil_div(n, 256);
  DBuf<uint64_t> bound(c, n);
  DBuf<uint32_t> fs(c, n), fmid(c, n), fl(c, n), fms(c, n), fm(c, n), fb8(c, n), fb(c, n), fd(c, n), caps(c, n);
  {
    DBuf<uint32_t> heavy(c, n), nheavy(c, 1);
    memset_async(c, nheavy.get(), 0, 4);
    HGP_LAUNCH(k_visit_bound, g, 256, 0, c, hoff, tch, toff, n, bound.get(), heavy.get(), nheavy.get());
    HGP_LAUNCH(k_visit_bound_heavy, std::min<uint32_t>(n, VISIT_HEAVY_GRID), 256, 0, c, hoff, tch, toff, heavy.get(),
               nheavy.get(), bound.get());
  }
  // the 8192-slot tier without shared-inbound counts (with them the big table already has 8192 slots)
  const uint32_t big8_cap = inter ? 0u : CAND_BIG8_SLOTS;
  HGP_LAUNCH(k_cand_tiers, g, 256, 0, c, bound.get(), n, direct_n, big8_cap, big_cap, fs.get(), fmid.get(), fl.get(),
             fms.get(), fm.get(), fb8.get(), fb.get(), fd.get(), caps.get());
  // tier lists: small, mid and light tables, small and large medium tables, 8192-slot tables, big or
  // direct, heavy
  const uint32_t* flags[TIERS - 1] = {fs.get(),  fmid.get(), fl.get(), fms.get(),
                                      fm.get(),  fb8.get(),  fb.get(), fd.get()};
  DBuf<uint64_t> pos[TIERS - 1];
  uint32_t cnt[TIERS];
  uint32_t rest = n;
  for (int i = 0; i < TIERS - 1; ++i) {
    pos[i] = DBuf<uint64_t>(c, (uint64_t)n + 1);
    exclusive_scan(c, flags[i], pos[i].get(), n);
  }
  for (int i = 0; i < TIERS - 1; ++i) {
    cnt[i] = (uint32_t)read_scalar(c, pos[i].get() + n);
    rest -= cnt[i];
  }
  cnt[TIERS - 1] = rest;
  DBuf<uint32_t> lists[TIERS];
  TierLists tl{};
  for (int i = 0; i < TIERS; ++i) {
    lists[i] = DBuf<uint32_t>(c, cnt[i]);
    tl.l[i] = lists[i].get();
    if (i < TIERS - 1) {
      tl.f[i] = flags[i];
      tl.p[i] = pos[i].get();
    }
  }
  HGP_LAUNCH(k_split_tiers, g, 256, 0, c, tl, n);
  const uint32_t ns = cnt[0], nmid = cnt[1], nl = cnt[2], nms = cnt[3], nm = cnt[4], nb8 = cnt[5], nb = cnt[6],
                 nd = cnt[7], nh = cnt[8];
  DBuf<uint32_t>&small = lists[0], &mid = lists[1], &light = lists[2], &med_s = lists[3], &med = lists[4],
                 &big8 = lists[5], &big = lists[6], &direct = lists[7], &heavy = lists[8];
  // the tiers write disjoint rows of pairs and scores: light and medium nodes on side streams
  // beside the big and heavy ones (which allocate, so they stay on the main stream)
  Fork fk(c, c.tune.side_streams & 1u);
  fk.on(0, [&] {
    // visit bounds up to CAND_LIGHT_S: 128-slot tables, 8 waves per workgroup
    if (inter)
      HGP_LAUNCH((k_cand_light<true, 2 * CAND_LIGHT_S, 8>), ceil_div(ns, 8), 512, 0, c, hedges, hoff, srcc, tch, toff,
                 inb, w, sizes, small.get(), ns, bound.get(), cnc, omega, delta, pairs, scores);
    else
      HGP_LAUNCH((k_cand_light<false, 2 * CAND_LIGHT_S, 8>), ceil_div(ns, 8), 512, 0, c, hedges, hoff, srcc, tch, toff,
                 inb, w, sizes, small.get(), ns, bound.get(), cnc, omega, delta, pairs, scores);
    if (inter)
      HGP_LAUNCH((k_cand_light<true, 2 * CAND_LIGHT_M, 8>), ceil_div(nmid, 8), 512, 0, c, hedges, hoff, srcc, tch,
                 toff, inb, w, sizes, mid.get(), nmid, bound.get(), cnc, omega, delta, pairs, scores);
    else
      HGP_LAUNCH((k_cand_light<false, 2 * CAND_LIGHT_M, 8>), ceil_div(nmid, 8), 512, 0, c, hedges, hoff, srcc, tch,
                 toff, inb, w, sizes, mid.get(), nmid, bound.get(), cnc, omega, delta, pairs, scores);
    if (inter)
      HGP_LAUNCH(k_cand_light<true>, ceil_div(nl, 4), 256, 0, c, hedges, hoff, srcc, tch, toff, inb, w, sizes,
                 light.get(), nl, bound.get(), cnc, omega, delta, pairs, scores);
    else
      HGP_LAUNCH(k_cand_light<false>, ceil_div(nl, 4), 256, 0, c, hedges, hoff, srcc, tch, toff, inb, w, sizes,
                 light.get(), nl, bound.get(), cnc, omega, delta, pairs, scores);
  });
#define CAND_LAUNCH(K, list, cnt, threads)                                                                          \
  HGP_LAUNCH(K, cnt, threads, 0, c, hedges, hoff, srcc, tch, toff, inb, w, sizes, list, caps.get(), cnc, omega, delta, \
             pairs, scores)
  fk.on(1, [&] {
    if (inter) CAND_LAUNCH((k_cand_medium<true, CAND_MED_S>), med_s.get(), nms, 256);
    else CAND_LAUNCH((k_cand_medium<false, CAND_MED_S>), med_s.get(), nms, 256);
    if (inter) CAND_LAUNCH((k_cand_medium<true, CAND_MED_CAP, 8>), med.get(), nm, 512);
    else CAND_LAUNCH((k_cand_medium<false, CAND_MED_CAP, 8>), med.get(), nm, 512);
  });
  if (nb8) CAND_LAUNCH(k_cand_big8, big8.get(), nb8, 1024);
  if (nb) {
    if (inter) CAND_LAUNCH(k_cand_big<true>, big.get(), nb, 1024);
    else CAND_LAUNCH(k_cand_big<false>, big.get(), nb, 1024);
  }
  if (nd) {  // only on levels of at most direct_n nodes
    if (inter) CAND_LAUNCH(k_cand_direct<true>, direct.get(), nd, 1024);
    else CAND_LAUNCH(k_cand_direct<false>, direct.get(), nd, 1024);
  }
#undef CAND_LAUNCH
  if (nh == 0) return;
  // heavy nodes: LDS attempt first, the overflow list goes to the HBM tables
  DBuf<uint32_t> ovf(c, nh), novf(c, 1);
  memset_async(c, novf.get(), 0, 4);
  if (inter)
    HGP_LAUNCH((k_cand_opt<true, 1>), nh, 1024, 0, c, hedges, hoff, srcc, tch, toff, inb, w, sizes, heavy.get(),
               caps.get(), cnc, omega, delta, pairs, scores, ovf.get(), novf.get());
  else
    HGP_LAUNCH((k_cand_opt<false, 1>), nh, 1024, 0, c, hedges, hoff, srcc, tch, toff, inb, w, sizes, heavy.get(),
               caps.get(), cnc, omega, delta, pairs, scores, ovf.get(), novf.get());
  uint32_t no = read_scalar(c, novf.get());
  const uint32_t no1 = no;
  if (no) {
    // second LDS attempt in OPT_CHUNKS passes; what still overflows goes to HBM tables
    DBuf<uint32_t> ovf_b(c, no);
    memset_async(c, novf.get(), 0, 4);
    if (inter)
      HGP_LAUNCH((k_cand_opt<true, OPT_CHUNKS>), no, 1024, 0, c, hedges, hoff, srcc, tch, toff, inb, w, sizes, ovf.get(),
                 caps.get(), cnc, omega, delta, pairs, scores, ovf_b.get(), novf.get());
    else
      HGP_LAUNCH((k_cand_opt<false, OPT_CHUNKS>), no, 1024, 0, c, hedges, hoff, srcc, tch, toff, inb, w, sizes,
                 ovf.get(), caps.get(), cnc, omega, delta, pairs, scores, ovf_b.get(), novf.get());
    no = read_scalar(c, novf.get());
    ovf = std::move(ovf_b);
  }
  if (c.verbose >= 2)
    fprintf(stderr, "  candidates: n=%u light %u medium %u big %u optimistic %u (overflow %u, then %u)\n", n, nl, nm,
            nb, nh, no1, no);
  if (no == 0) return;
  // HBM tables: first with tables capped at c.tune.heavy_slots (filled to heavy_fill16 / 16, the
  // autotuner's choices; 65,536 and 3/4 by default), then the overflow with full ones
  DBuf<uint32_t> ovf2(c, no), novf2(c, 1), mc(c, 1);
  const uint32_t* lst = ovf.get();
  uint32_t cntl = no;
  for (int pass = 0; pass < 2 && cntl; ++pass) {
    const uint32_t try_cap = pass == 0 ? c.tune.heavy_slots : 0u;
    const uint32_t try_fill = (uint32_t)((uint64_t)try_cap * c.tune.heavy_fill16 / 16);
    memset_async(c, mc.get(), 0, 4);
    memset_async(c, novf2.get(), 0, 4);
    HGP_LAUNCH(k_heavy_max_cap, std::min<uint32_t>(ceil_div(cntl, 256), 1024), 256, 0, c, lst, caps.get(), cntl,
               try_cap ? try_cap : 0xFFFFFFFFu, mc.get());
    const uint64_t region = read_scalar(c, mc.get());
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>({(uint64_t)cntl, HEAVY_TABLE_SLOTS / region, 4096}));
    DBuf<uint32_t> gk(c, grid * region), gs(c, grid * region), gi(c, grid * region);
    HGP_LAUNCH(k_cand_heavy, (uint32_t)grid, 256, 0, c, hedges, hoff, srcc, tch, toff, inb, w, sizes, lst, caps.get(),
               cnc, omega, delta, pairs, scores, cntl, n, region, gk.get(), gs.get(), gi.get(), try_cap, try_fill, ovf2.get(),
               novf2.get());
    if (c.verbose >= 2)
      fprintf(stderr, "  candidates: %u nodes in HBM tables (pass %d, region %llu, grid %llu)\n", cntl, pass,
              (unsigned long long)region, (unsigned long long)grid);
    if (pass == 0) {
      cntl = read_scalar(c, novf2.get());
      lst = ovf2.get();
    }
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
