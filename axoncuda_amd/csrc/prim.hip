// prim.hip - scans, stable radix sort and small helpers for gfx950 (wave64).
//
// Scans: one pass with decoupled look-back over 4096-element tiles.
// Radix sort: reduce-then-scan LSD with 8-bit digits; the scatter ranks keys within a tile
// in 256-key rounds, using eight ballots per key to find same-digit lanes of a wave (a
// stable in-wave rank) and LDS prefix sums across the four waves of the workgroup.
#include "prim.hpp"

namespace hgp {

namespace {

// Single-pass scan with decoupled look-back.  A workgroup takes the next tile number from a
// ticket counter (so every tile it waits for belongs to a workgroup that is already running),
// scans its 4096 elements, publishes its aggregate, then walks back over the predecessors'
// published aggregates / inclusive prefixes until it meets an inclusive one.  Flags carry a
// per-scan generation, so the status arrays are reused without clearing.
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

template <typename TI, typename TO, bool EXCL>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_lookback(const TI* in, TO* out, uint64_t n, uint32_t tiles,
                                                              uint32_t* flags, TO* agg, TO* incl,
                                                              unsigned long long* ticket, uint64_t base,
                                                              uint32_t gen) {
  __shared__ uint32_t s_tile;
  __shared__ TO s_wave[SCAN_BLOCK / 64];
  __shared__ TO s_prefix;
  if (threadIdx.x == 0) s_tile = (uint32_t)(atomicAdd(ticket, 1ull) - base);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t w = threadIdx.x / 64, l = lane_id();
  const uint64_t start = (uint64_t)tile * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  TO v[SCAN_ITEMS];
  TO sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint64_t i = start + k;
    v[k] = i < n ? (TO)in[i] : (TO)0;
    sum += v[k];
  }
  const TO winc = wave_scan_incl(sum);
  if (l == 63) s_wave[w] = winc;
  __syncthreads();
  TO wbase = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < SCAN_BLOCK / 64; ++i) {
    const TO x = s_wave[i];
    if (i < (int)w) wbase += x;
    tot += x;
  }
  const TO texcl = wbase + winc - sum;
  // look-back by wave 0: each round examines the 64 preceding tiles at once
  if (w == 0) {
    const uint32_t g = gen << 2;
    TO prefix = 0;
    if (tile == 0) {
      if (l == 0) {
        incl[0] = tot;
        __hip_atomic_store(&flags[0], g | 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    } else {
      if (l == 0) {
        agg[tile] = tot;
        __hip_atomic_store(&flags[tile], g | 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
      int64_t hi = (int64_t)tile - 1;  // window [hi - 63, hi]; lane l looks at tile hi - l
      for (;;) {
        const int64_t j = hi - (int64_t)l;
        uint32_t st = 2u;  // tiles before 0 act as an inclusive zero
        if (j >= 0) {
          const uint32_t f = __hip_atomic_load(&flags[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          st = (f & ~3u) == g ? (f & 3u) : 0u;
        }
        const uint64_t incl_mask = __ballot(st == 2u);
        const uint64_t zero_mask = __ballot(st == 0u);
        // nearest inclusive tile (lowest lane with st == 2); every lane before it must be published
        const uint32_t first_incl = incl_mask ? (uint32_t)__ffsll((unsigned long long)incl_mask) - 1u : 64u;
        const uint64_t before = first_incl >= 64 ? ~0ull : ((1ull << first_incl) - 1ull);
        if (zero_mask & before) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        TO part = 0;
        if (l < first_incl && j >= 0) part = agg[j];
        if (l == first_incl && j >= 0) part = incl[j];
        prefix += wave_sum(part);
        if (first_incl < 64) break;
        hi -= 64;
      }
      if (l == 0) {
        incl[tile] = prefix + tot;
        __hip_atomic_store(&flags[tile], g | 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (l == 0) {
      s_prefix = prefix;
      if (EXCL && tile == tiles - 1) out[n] = prefix + tot;
    }
  }
  __syncthreads();
  TO run = s_prefix + texcl;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint64_t i = start + k;
    if (EXCL) {
      if (i < n) out[i] = run;
      run += v[k];
    } else {
      run += v[k];
      if (i < n) out[i] = run;
    }
  }
}

// Three-pass scan for large inputs (tile reduce, scan of tile sums, tile scan): at thousands
// of tiles the decoupled look-back's chain of publications is slower than two extra passes.
constexpr int RS3_ITEMS = 8;
constexpr int RS3_TILE = SCAN_BLOCK * RS3_ITEMS;

template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* lds_waves, T* total) {
  // lds_waves: SCAN_BLOCK/64 entries
  const uint32_t w = threadIdx.x / 64, l = lane_id();
  T incl = wave_scan_incl(v);
  if (l == 63) lds_waves[w] = incl;
  __syncthreads();
  T base = 0, tot = 0;
  for (uint32_t i = 0; i < blockDim.x / 64; ++i) {
    T x = lds_waves[i];
    if (i < w) base += x;
    tot += x;
  }
  *total = tot;
  __syncthreads();
  return base + incl - v;
}

template <typename TI, typename TO>
__global__ void k_tile_reduce(const TI* __restrict__ in, uint64_t n, TO* __restrict__ partial) {
  __shared__ TO sw[SCAN_BLOCK / 64];
  uint64_t base = (uint64_t)blockIdx.x * RS3_TILE;
  TO s = 0;
  for (int k = 0; k < RS3_ITEMS; ++k) {
    uint64_t i = base + (uint64_t)k * SCAN_BLOCK + threadIdx.x;
    if (i < n) s += (TO)in[i];
  }
  s = wave_sum(s);
  if (lane_id() == 0) sw[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    TO t = 0;
    for (int i = 0; i < SCAN_BLOCK / 64; ++i) t += sw[i];
    partial[blockIdx.x] = t;
  }
}

// exclusive scan of partial[0..m) in place by one workgroup; partial[m] = total
template <typename TO>
__global__ void k_scan_partials(TO* partial, uint32_t m) {
  __shared__ TO sw[1024 / 64];
  __shared__ TO carry_s;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  for (uint32_t base = 0; base < m; base += blockDim.x) {
    uint32_t i = base + threadIdx.x;
    TO v = i < m ? partial[i] : 0;
    TO tot;
    TO ex = block_excl_scan(v, sw, &tot);
    TO carry = carry_s;
    if (i < m) partial[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry_s = carry + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[m] = carry_s;
}

// Also stores the total at out[n] (block 0): the partials' scan left it at partial[gridDim.x].
template <typename TI, typename TO>
__global__ void k_tile_scan_excl(const TI* in, TO* out, uint64_t n, const TO* __restrict__ partial) {
  __shared__ TO sw[SCAN_BLOCK / 64];
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = partial[gridDim.x];
  uint64_t base = (uint64_t)blockIdx.x * RS3_TILE + (uint64_t)threadIdx.x * RS3_ITEMS;
  TO v[RS3_ITEMS];
  TO s = 0;
#pragma unroll
  for (int k = 0; k < RS3_ITEMS; ++k) {
    uint64_t i = base + k;
    v[k] = i < n ? (TO)in[i] : (TO)0;
    s += v[k];
  }
  TO tot;
  TO ex = block_excl_scan(s, sw, &tot) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < RS3_ITEMS; ++k) {
    uint64_t i = base + k;
    if (i < n) out[i] = ex;
    ex += v[k];
  }
}

template <typename TI, typename TO>
__global__ void k_tile_scan_incl(const TI* in, TO* out, uint64_t n, const TO* __restrict__ partial) {
  __shared__ TO sw[SCAN_BLOCK / 64];
  uint64_t base = (uint64_t)blockIdx.x * RS3_TILE + (uint64_t)threadIdx.x * RS3_ITEMS;
  TO v[RS3_ITEMS];
  TO s = 0;
#pragma unroll
  for (int k = 0; k < RS3_ITEMS; ++k) {
    uint64_t i = base + k;
    v[k] = i < n ? (TO)in[i] : (TO)0;
    s += v[k];
  }
  TO tot;
  TO ex = block_excl_scan(s, sw, &tot) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < RS3_ITEMS; ++k) {
    uint64_t i = base + k;
    ex += v[k];
    if (i < n) out[i] = ex;
  }
}


// Status arrays of the context's scans (one stream: scans never overlap).
void scan_status(Ctx& c, uint32_t tiles) {
  if (tiles <= c.scan_cap_tiles && c.scan_ticket) return;
  uint32_t cap = 1024;
  while (cap < tiles) cap <<= 1;
  if (c.scan_flags) {
    dfree(c, c.scan_flags);
    dfree(c, c.scan_agg);
    dfree(c, c.scan_incl);
  }
  c.scan_flags = (uint32_t*)dmalloc(c, (uint64_t)cap * 4);
  c.scan_agg = dmalloc(c, (uint64_t)cap * 8);
  c.scan_incl = dmalloc(c, (uint64_t)cap * 8);
  memset_async(c, c.scan_flags, 0, (uint64_t)cap * 4);  // generation 0 is never used
  if (!c.scan_ticket) {
    c.scan_ticket = (unsigned long long*)dmalloc(c, 8);
    memset_async(c, c.scan_ticket, 0, 8);
    c.scan_tickets_issued = 0;
  }
  c.scan_cap_tiles = cap;
}

// Grid-barrier counter of the context's cooperative sorts: returns the count that a launch
// making `arrivals` arrivals starts from, resetting the counter (on the stream) before it
// could wrap.  The caller adds the arrivals to c.coop_issued once the launch is queued, so a
// launch that fails to queue leaves the count as the device has it.
unsigned int coop_barrier_base(Ctx& c, uint32_t arrivals) {
  if (!c.coop_bar) {
    c.coop_bar = (unsigned int*)dmalloc(c, 4);
    c.coop_issued = 1u << 31;  // forces the reset below
  }
  if (c.coop_issued + (uint64_t)arrivals >= (1ull << 31)) {
    fill_u32(c, c.coop_bar, 0u, 1);
    c.coop_issued = 0;
  }
  return (unsigned int)c.coop_issued;
}

// in == out is safe: each workgroup reads only its own tile before writing it.
template <typename TI, typename TO>
void scan_impl(Ctx& c, const TI* in, TO* out, uint64_t n, bool exclusive) {
  if (n == 0) {
    if (exclusive) memset_async(c, out, 0, sizeof(TO));
    return;
  }
  const uint32_t tiles = ceil_div(n, SCAN_TILE);
  if (tiles > c.tune.scan_lookback_tiles) {  // default 64 (the autotuner may move it)
    const uint32_t t3 = ceil_div(n, RS3_TILE);
    DBuf<TO> partial(c, (uint64_t)t3 + 1);
    HGP_LAUNCH((k_tile_reduce<TI, TO>), t3, SCAN_BLOCK, 0, c, in, n, partial.get());
    HGP_LAUNCH((k_scan_partials<TO>), 1, 1024, 0, c, partial.get(), t3);
    if (exclusive) {
      HGP_LAUNCH((k_tile_scan_excl<TI, TO>), t3, SCAN_BLOCK, 0, c, in, out, n, partial.get());
    } else {
      HGP_LAUNCH((k_tile_scan_incl<TI, TO>), t3, SCAN_BLOCK, 0, c, in, out, n, partial.get());
    }
    return;
  }
  scan_status(c, tiles);
  c.scan_gen = (c.scan_gen + 1) & 0x3FFFFFFFu;
  if (c.scan_gen == 0) c.scan_gen = 1;
  const uint64_t base = c.scan_tickets_issued;
  c.scan_tickets_issued += tiles;
  if (exclusive)
    HGP_LAUNCH((k_scan_lookback<TI, TO, true>), tiles, SCAN_BLOCK, 0, c, in, out, n, tiles, c.scan_flags,
               (TO*)c.scan_agg, (TO*)c.scan_incl, c.scan_ticket, base, c.scan_gen);
  else
    HGP_LAUNCH((k_scan_lookback<TI, TO, false>), tiles, SCAN_BLOCK, 0, c, in, out, n, tiles, c.scan_flags,
               (TO*)c.scan_agg, (TO*)c.scan_incl, c.scan_ticket, base, c.scan_gen);
}

// ---------------------------------------------------------------------------------------
// radix sort
constexpr int RS_BLOCK = 256;
constexpr uint32_t RS_MAX_ROUNDS = 16;  // keys per thread and tile of a large sort
constexpr uint32_t RS_TARGET_TILES = 512;  // small sorts shrink their tiles to spread over the CUs

// Rounds of RS_BLOCK keys per tile: one tile per workgroup, so a small sort runs many short
// workgroups side by side instead of a few that walk 16 rounds one after another.  The
// autotuner may set the rounds per size class (floor(log2 n)); the result is the same.
inline uint32_t rs_rounds(const Ctx& c, uint64_t n) {
  const int cls = bit_width(n) - 1;
  if (cls >= 0 && cls < 40 && c.tune.sort_rounds[cls]) return std::min(c.tune.sort_rounds[cls], RS_MAX_ROUNDS);
  return (uint32_t)std::min<uint64_t>(RS_MAX_ROUNDS, std::max<uint64_t>(1, ceil_div(n, (uint64_t)RS_BLOCK * RS_TARGET_TILES)));
}

template <typename K>
__global__ void k_rs_hist(const K* __restrict__ keys, uint64_t n, int shift, uint32_t nblocks, uint32_t rounds,
                          K kmin, uint32_t* __restrict__ counts) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint64_t base = (uint64_t)blockIdx.x * RS_BLOCK * rounds;
  for (uint32_t r = 0; r < rounds; ++r) {
    uint64_t i = base + (uint64_t)r * RS_BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)((keys[i] - kmin) >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  counts[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

template <typename K, bool HAS_VALS>
__global__ void __launch_bounds__(RS_BLOCK) k_rs_scatter(const K* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                         K* __restrict__ okeys, uint32_t* __restrict__ ovals,
                                                         uint64_t n, int shift, uint32_t nblocks, uint32_t rounds,
                                                         K kmin, const uint64_t* __restrict__ offsets) {
  __shared__ uint64_t run[256];
  __shared__ uint32_t wcnt[RS_BLOCK / 64][256];
  const uint32_t t = threadIdx.x, w = t / 64;
  run[t] = offsets[(uint64_t)t * nblocks + blockIdx.x];
  uint64_t base = (uint64_t)blockIdx.x * RS_BLOCK * rounds;
  for (uint32_t r = 0; r < rounds; ++r) {
    uint64_t i = base + (uint64_t)r * RS_BLOCK + t;
    if (base + (uint64_t)r * RS_BLOCK >= n) break;  // uniform across the block
#pragma unroll
    for (int q = 0; q < RS_BLOCK / 64; ++q) wcnt[q][t] = 0;
    __syncthreads();
    const bool valid = i < n;
    K key = valid ? keys[i] : (K)0;
    uint32_t val = 0;
    if (HAS_VALS && valid) val = vals[i];
    const uint32_t d = (uint32_t)((key - kmin) >> shift) & 0xFFu;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = lanes_below(peers);
    if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    // per digit: exclusive prefix over waves (thread t handles digit t)
    uint32_t acc = 0;
#pragma unroll
    for (int q = 0; q < RS_BLOCK / 64; ++q) {
      uint32_t c = wcnt[q][t];
      wcnt[q][t] = acc;
      acc += c;
    }
    __syncthreads();
    if (valid) {
      uint64_t pos = run[d] + wcnt[w][d] + rank;
      okeys[pos] = key;
      if (HAS_VALS) ovals[pos] = val;
    }
    __syncthreads();
    run[t] += acc;
  }
}

// Whole LSD sort of a small array by one workgroup: each pass (digit histogram, digit
// offsets, stable scatter in rounds of RS1_BLOCK keys) runs between barriers, so a small sort
// costs one launch instead of three or four per pass.  Keys of the pass stay in registers
// between the histogram and the scatter.  Ends in kb/vb after an odd pass count.
constexpr int RS1_BLOCK = 1024;
constexpr uint32_t RS1_MAX_ROUNDS = 16;
constexpr uint64_t RS1_MAX_N = (uint64_t)RS1_BLOCK * RS1_MAX_ROUNDS;

template <typename K, bool HAS_VALS>
__global__ void __launch_bounds__(RS1_BLOCK) k_rs_small(K* ka, uint32_t* va, K* kb, uint32_t* vb, uint32_t n,
                                                        int lo, int bits, K kmin) {
  __shared__ uint32_t h[256];
  __shared__ uint32_t run[256];
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t wcnt[RS1_BLOCK / 64][256];
  const uint32_t t = threadIdx.x, w = t / 64, l = lane_id();
  const uint32_t rounds = (n + RS1_BLOCK - 1) / RS1_BLOCK;
  for (int shift = lo; shift < bits; shift += 8) {
    K key[RS1_MAX_ROUNDS];
    uint32_t val[RS1_MAX_ROUNDS];
    if (t < 256) h[t] = 0;
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < RS1_MAX_ROUNDS; ++r) {
      const uint32_t i = r * RS1_BLOCK + t;
      key[r] = 0;
      val[r] = 0;
      if (i < n) {
        key[r] = ka[i];
        if (HAS_VALS) val[r] = va[i];
        atomicAdd(&h[(uint32_t)((key[r] - kmin) >> shift) & 0xFFu], 1u);
      }
    }
    __syncthreads();
    // run[d]: keys of digits below d (exclusive scan of the histogram by the first four waves)
    uint32_t x = 0, inc = 0;
    if (t < 256) {
      x = h[t];
      inc = wave_scan_incl(x);
      if (l == 63) s_w[w] = inc;
    }
    __syncthreads();
    if (t < 256) {
      uint32_t b = 0;
      for (uint32_t q = 0; q < w; ++q) b += s_w[q];
      run[t] = b + inc - x;
    }
#pragma unroll
    for (uint32_t r = 0; r < RS1_MAX_ROUNDS; ++r) {
      if (r < rounds) {  // uniform across the block
#pragma unroll
        for (int q = 0; q < 4; ++q) (&wcnt[0][0])[q * RS1_BLOCK + t] = 0;
        __syncthreads();
        const uint32_t i = r * RS1_BLOCK + t;
        const bool valid = i < n;
        const uint32_t d = (uint32_t)((key[r] - kmin) >> shift) & 0xFFu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const bool bit = (d >> b) & 1u;
          const uint64_t bb = __ballot(bit);
          peers &= bit ? bb : ~bb;
        }
        const uint32_t rank = lanes_below(peers);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        // per digit (thread t < 256 handles digit t): exclusive prefix over the waves
        uint32_t acc = 0;
        if (t < 256) {
#pragma unroll
          for (int q = 0; q < RS1_BLOCK / 64; ++q) {
            const uint32_t cq = wcnt[q][t];
            wcnt[q][t] = acc;
            acc += cq;
          }
        }
        __syncthreads();
        if (valid) {
          const uint32_t pos = run[d] + wcnt[w][d] + rank;
          kb[pos] = key[r];
          if (HAS_VALS) vb[pos] = val[r];
        }
        __syncthreads();
        if (t < 256) run[t] += acc;
      }
    }
    K* tk = ka;
    ka = kb;
    kb = tk;
    uint32_t* tv = va;
    va = vb;
    vb = tv;
    __threadfence();  // the next pass reads what this one wrote
    __syncthreads();
  }
}

// Whole LSD sort of a mid-size array in one launch: B workgroups each own a contiguous tile and
// meet at grid barriers.  B <= COOP_MAX_BLOCKS = 64 is far below what the device holds at once
// (at least 4 such workgroups per CU, 256 CUs), and every other kernel that may run beside one
// finishes on its own (look-back scans wait only on tiles already running; serving workers do
// not use these sorts), so all B workgroups become resident and every barrier is reached.  A pass counts its tile's digits, publishes the counts, and after a
// barrier every workgroup derives its tile's digit offsets from all counts (digit totals
// scanned over digits, plus the counts of the tiles before its own), scatters its keys stably
// in rounds of 256, and waits at a second barrier before the next pass reads the output.
// The barrier counter only grows; arrival k of a launch waits for base + B * k.
constexpr uint32_t COOP_MAX_BLOCKS = 64;

__device__ __forceinline__ void grid_barrier(unsigned int* ctr, unsigned int target) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this thread's writes before the arrival
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(ctr, 1u);
    while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) __builtin_amdgcn_s_sleep(1);
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the other workgroups' writes after it
}

template <typename K, bool HAS_VALS>
__global__ void __launch_bounds__(RS_BLOCK) k_rs_coop(K* ka, uint32_t* va, K* kb, uint32_t* vb, uint32_t n,
                                                      uint32_t tile, int lo, int bits, K kmin,
                                                      uint32_t* __restrict__ counts, unsigned int* bar,
                                                      unsigned int bar_base) {
  __shared__ uint32_t h[256];
  __shared__ uint32_t run[256];
  __shared__ uint32_t s_w[RS_BLOCK / 64];
  __shared__ uint32_t wcnt[RS_BLOCK / 64][256];
  const uint32_t B = gridDim.x, b = blockIdx.x, t = threadIdx.x, w = t / 64, l = lane_id();
  const uint32_t beg = min(n, b * tile), end = min(n, beg + tile);
  unsigned int target = bar_base;
  for (int shift = lo; shift < bits; shift += 8) {
    h[t] = 0;
    __syncthreads();
    for (uint32_t i = beg + t; i < end; i += RS_BLOCK) atomicAdd(&h[(uint32_t)((ka[i] - kmin) >> shift) & 0xFFu], 1u);
    __syncthreads();
    counts[(uint64_t)b * 256 + t] = h[t];
    target += B;
    grid_barrier(bar, target);
    // digit t: its total over all tiles and its count in the tiles before this one
    uint32_t tot = 0, pre = 0;
    for (uint32_t x = 0; x < B; ++x) {
      const uint32_t cx = counts[(uint64_t)x * 256 + t];
      tot += cx;
      pre += x < b ? cx : 0u;
    }
    const uint32_t inc = wave_scan_incl(tot);
    if (l == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t q = 0; q < w; ++q) base += s_w[q];
    run[t] = base + inc - tot + pre;
    for (uint32_t r0 = beg; r0 < end; r0 += RS_BLOCK) {  // uniform across the workgroup
#pragma unroll
      for (int q = 0; q < RS_BLOCK / 64; ++q) wcnt[q][t] = 0;
      __syncthreads();
      const uint32_t i = r0 + t;
      const bool valid = i < end;
      const K key = valid ? ka[i] : (K)0;
      uint32_t val = 0;
      if (HAS_VALS && valid) val = va[i];
      const uint32_t d = (uint32_t)((key - kmin) >> shift) & 0xFFu;
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int bt = 0; bt < 8; ++bt) {
        const bool bit = (d >> bt) & 1u;
        const uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
      }
      const uint32_t rank = lanes_below(peers);
      if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
      __syncthreads();
      uint32_t acc = 0;
#pragma unroll
      for (int q = 0; q < RS_BLOCK / 64; ++q) {
        const uint32_t cq = wcnt[q][t];
        wcnt[q][t] = acc;
        acc += cq;
      }
      __syncthreads();
      if (valid) {
        const uint32_t pos = run[d] + wcnt[w][d] + rank;
        kb[pos] = key;
        if (HAS_VALS) vb[pos] = val;
      }
      __syncthreads();
      run[t] += acc;
    }
    target += B;
    grid_barrier(bar, target);  // the output is whole (and the counts free) before the next pass
    K* tk = ka;
    ka = kb;
    kb = tk;
    uint32_t* tv = va;
    va = vb;
    vb = tv;
  }
}

// LSD passes over bits [lo, bits) of key - kmin (the keys themselves are moved unchanged).  With `kbuf` (and `vbuf`) the result ends up in those
// buffers by exchanging them with the scratch ones after an odd pass count; otherwise it is
// copied back into keys / vals.
template <typename K>
void radix_impl(Ctx& c, K* keys, uint32_t* vals, uint64_t n, int bits, int lo = 0, DBuf<K>* kbuf = nullptr,
                DBuf<uint32_t>* vbuf = nullptr, K kmin = 0) {
  if (n <= 1 || bits <= lo) return;
  if (n <= RS1_MAX_N) {
    DBuf<K> k2(c, n);
    DBuf<uint32_t> v2(c, vals ? n : 0);
    const int passes = (bits - lo + 7) / 8;
    // bytes: keys (and values) read once and written once per pass
    const uint64_t sb = (uint64_t)passes * n * (2 * sizeof(K) + (vals ? 8 : 0));
    if (vals)
      HGP_LAUNCH_B(sb, (k_rs_small<K, true>), 1, RS1_BLOCK, 0, c, keys, vals, k2.get(), v2.get(), (uint32_t)n, lo,
                   bits, kmin);
    else
      HGP_LAUNCH_B(sb, (k_rs_small<K, false>), 1, RS1_BLOCK, 0, c, keys, vals, k2.get(), v2.get(), (uint32_t)n, lo,
                   bits, kmin);
    if (passes & 1) {
      if (kbuf && (!vals || vbuf)) {  // hand the scratch buffers over instead of copying
        std::swap(*kbuf, k2);
        if (vals) std::swap(*vbuf, v2);
        return;
      }
      copy_d2d(c, keys, k2.get(), n * sizeof(K));
      if (vals) copy_d2d(c, vals, v2.get(), n * sizeof(uint32_t));
    }
    return;
  }
  if (n <= c.tune.coop_max_n) {
    // one launch for every pass: ceil(n / coop_tile) workgroups (at most COOP_MAX_BLOCKS)
    const uint32_t tile0 = std::max<uint32_t>(c.tune.coop_tile, (uint32_t)ceil_div(n, COOP_MAX_BLOCKS));
    const uint32_t tile = (tile0 + RS_BLOCK - 1) / RS_BLOCK * RS_BLOCK;
    const uint32_t B = ceil_div(n, tile);
    const int passes = (bits - lo + 7) / 8;
    const unsigned int base = coop_barrier_base(c, 2u * passes * B);
    DBuf<K> k2(c, n);
    DBuf<uint32_t> v2(c, vals ? n : 0);
    DBuf<uint32_t> counts(c, (uint64_t)256 * B);
    const uint64_t sb = (uint64_t)passes * n * (2 * sizeof(K) + (vals ? 8 : 0));
    if (vals)
      HGP_LAUNCH_B(sb, (k_rs_coop<K, true>), B, RS_BLOCK, 0, c, keys, vals, k2.get(), v2.get(), (uint32_t)n, tile, lo,
                   bits, kmin, counts.get(), c.coop_bar, base);
    else
      HGP_LAUNCH_B(sb, (k_rs_coop<K, false>), B, RS_BLOCK, 0, c, keys, vals, k2.get(), v2.get(), (uint32_t)n, tile,
                   lo, bits, kmin, counts.get(), c.coop_bar, base);
    c.coop_issued += 2u * passes * B;
    if (passes & 1) {
      if (kbuf && (!vals || vbuf)) {  // hand the scratch buffers over instead of copying
        std::swap(*kbuf, k2);
        if (vals) std::swap(*vbuf, v2);
        return;
      }
      copy_d2d(c, keys, k2.get(), n * sizeof(K));
      if (vals) copy_d2d(c, vals, v2.get(), n * sizeof(uint32_t));
    }
    return;
  }
  const uint32_t rounds = rs_rounds(c, n);
  const uint32_t nblocks = ceil_div(n, (uint64_t)RS_BLOCK * rounds);
  DBuf<K> k2(c, n);
  DBuf<uint32_t> v2(c, vals ? n : 0);
  DBuf<uint32_t> counts(c, (uint64_t)256 * nblocks);
  DBuf<uint64_t> offsets(c, (uint64_t)256 * nblocks + 1);
  K* ka = keys;
  K* kb = k2.get();
  uint32_t* va = vals;
  uint32_t* vb = v2.get();
  for (int shift = lo; shift < bits; shift += 8) {
    // bytes: keys read once, one count per (digit, tile) written
    HGP_LAUNCH_B(n * sizeof(K) + 1024ull * nblocks, (k_rs_hist<K>), nblocks, RS_BLOCK, 0, c, ka, n, shift, nblocks, rounds, kmin,
                 counts.get());
    exclusive_scan(c, counts.get(), offsets.get(), (uint64_t)256 * nblocks);
    // bytes: keys (and values) read once and written once, one offset per (digit, tile) read
    const uint64_t sb = n * (2 * sizeof(K) + (vals ? 8 : 0)) + 2048ull * nblocks;
    if (vals)
      HGP_LAUNCH_B(sb, (k_rs_scatter<K, true>), nblocks, RS_BLOCK, 0, c, ka, va, kb, vb, n, shift, nblocks,
                   rounds, kmin, offsets.get());
    else
      HGP_LAUNCH_B(sb, (k_rs_scatter<K, false>), nblocks, RS_BLOCK, 0, c, ka, va, kb, vb, n, shift, nblocks,
                   rounds, kmin, offsets.get());
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != keys) {
    if (kbuf && (!vals || vbuf)) {  // hand the scratch buffers over instead of copying
      std::swap(*kbuf, k2);
      if (vals) std::swap(*vbuf, v2);
      return;
    }
    copy_d2d(c, keys, ka, n * sizeof(K));
    if (vals) copy_d2d(c, vals, va, n * sizeof(uint32_t));
  }
}

__global__ void k_minmax_u64(const uint64_t* __restrict__ p, uint64_t n, unsigned long long* out) {
  uint64_t mn = ~0ull, mx = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = p[i];
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
  }
  __shared__ uint64_t s_mn[4], s_mx[4];
  mn = wave_min(mn);
  mx = wave_max(mx);
  if (lane_id() == 0) {
    s_mn[threadIdx.x / 64] = mn;
    s_mx[threadIdx.x / 64] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 4; ++i) {
      mn = s_mn[i] < mn ? s_mn[i] : mn;
      mx = s_mx[i] > mx ? s_mx[i] : mx;
    }
    atomicMin(&out[0], (unsigned long long)mn);
    atomicMax(&out[1], (unsigned long long)mx);
  }
}
__global__ void k_init_minmax(unsigned long long* mm) {
  mm[0] = ~0ull;
  mm[1] = 0ull;
}

template <typename K>
__global__ void k_segment_offsets(const K* __restrict__ keys, int shift, uint64_t n, uint32_t nseg,
                                  uint64_t* __restrict__ off) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  int64_t a = i == 0 ? -1 : (int64_t)(uint32_t)(keys[i - 1] >> shift);
  int64_t b = i == n ? (int64_t)nseg : (int64_t)(uint32_t)(keys[i] >> shift);
  if (b > (int64_t)nseg) b = nseg;
  for (int64_t s = a + 1; s <= b; ++s) off[s] = i;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_iota_u32(uint32_t* p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}
__global__ void k_max_u64(const uint64_t* __restrict__ p, uint64_t n, unsigned long long* out) {
  uint64_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    m = p[i] > m ? p[i] : m;
  m = wave_max(m);
  if (lane_id() == 0) atomicMax(out, (unsigned long long)m);
}

__global__ void k_max_u32(const uint32_t* __restrict__ p, uint64_t n, uint32_t* out) {
  uint32_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    m = p[i] > m ? p[i] : m;
  m = wave_max(m);
  if (lane_id() == 0) atomicMax(out, m);
}

}  // namespace

void exclusive_scan(Ctx& c, const uint32_t* in, uint64_t* out, uint64_t n) { scan_impl<uint32_t, uint64_t>(c, in, out, n, true); }
void exclusive_scan(Ctx& c, const uint64_t* in, uint64_t* out, uint64_t n) { scan_impl<uint64_t, uint64_t>(c, in, out, n, true); }
void inclusive_scan(Ctx& c, const int64_t* in, int64_t* out, uint64_t n) { scan_impl<int64_t, int64_t>(c, in, out, n, false); }
void inclusive_scan(Ctx& c, const int32_t* in, int32_t* out, uint64_t n) { scan_impl<int32_t, int32_t>(c, in, out, n, false); }

void radix_sort(Ctx& c, uint32_t* keys, uint32_t* vals, uint64_t n, int bits) { radix_impl<uint32_t>(c, keys, vals, n, bits); }
void radix_sort(Ctx& c, uint64_t* keys, uint32_t* vals, uint64_t n, int bits) { radix_impl<uint64_t>(c, keys, vals, n, bits); }
void radix_sort_bits(Ctx& c, uint64_t* keys, uint32_t* vals, uint64_t n, int lo, int hi) {
  radix_impl<uint64_t>(c, keys, vals, n, hi, lo);
}
void radix_sort_buf(Ctx& c, DBuf<uint32_t>& keys, DBuf<uint32_t>& vals, uint64_t n, int lo, int hi) {
  radix_impl<uint32_t>(c, keys.get(), vals.get(), n, hi, lo, &keys, &vals);
}
void radix_sort_buf(Ctx& c, DBuf<uint64_t>& keys, DBuf<uint32_t>& vals, uint64_t n, int lo, int hi) {
  radix_impl<uint64_t>(c, keys.get(), vals.get(), n, hi, lo, &keys, &vals);
}

void segment_offsets(Ctx& c, const uint32_t* keys, int shift, uint64_t n, uint32_t nseg, uint64_t* off) {
  HGP_LAUNCH((k_segment_offsets<uint32_t>), ceil_div(n + 1, 256), 256, 0, c, keys, shift, n, nseg, off);
}
void segment_offsets(Ctx& c, const uint64_t* keys, int shift, uint64_t n, uint32_t nseg, uint64_t* off) {
  HGP_LAUNCH((k_segment_offsets<uint64_t>), ceil_div(n + 1, 256), 256, 0, c, keys, shift, n, nseg, off);
}

namespace {
void range_impl(Ctx& c, uint64_t* keys, uint32_t* vals, uint64_t n, DBuf<uint64_t>* kbuf, DBuf<uint32_t>* vbuf) {
  if (n <= 1) return;
  DBuf<unsigned long long> mm(c, 2);
  HGP_LAUNCH(k_init_minmax, 1, 1, 0, c, mm.get());  // on the stream: no host-to-device copy
  HGP_LAUNCH(k_minmax_u64, std::min<uint32_t>(ceil_div(n, 256), 512), 256, 0, c, keys, n, mm.get());
  unsigned long long h[2];
  copy_d2h(c, h, mm.get(), sizeof h);
  if (h[1] == h[0]) return;  // all keys equal: a stable sort is the identity
  // digits of key - min: the passes cover only the keys' range
  radix_impl<uint64_t>(c, keys, vals, n, bit_width(h[1] - h[0]), 0, kbuf, vbuf, (uint64_t)h[0]);
}
}  // namespace

void radix_sort_range(Ctx& c, uint64_t* keys, uint32_t* vals, uint64_t n) { range_impl(c, keys, vals, n, nullptr, nullptr); }
void radix_sort_range_buf(Ctx& c, DBuf<uint64_t>& keys, DBuf<uint32_t>& vals, uint64_t n) {
  range_impl(c, keys.get(), vals.get(), n, &keys, &vals);
}

Fork::Fork(Ctx& c, bool enabled) : c_(c), enabled_(enabled) {
  if (!enabled_) return;
  if (!c_.fork_ev) HGP_HIP(hipEventCreateWithFlags(&c_.fork_ev, hipEventDisableTiming));
  HGP_HIP(hipEventRecord(c_.fork_ev, c_.stream));
}
void Fork::start(int i) {
  HGP_REQUIRE(i >= 0 && i < Ctx::SIDES, HGP_E_INVALID, "Fork: no such side stream");
  if (used_ & (1u << i)) return;
  if (!c_.side[i]) {
    HGP_HIP(hipStreamCreateWithFlags(&c_.side[i], hipStreamNonBlocking));
    HGP_HIP(hipEventCreateWithFlags(&c_.side_ev[i], hipEventDisableTiming));
  }
  HGP_HIP(hipStreamWaitEvent(c_.side[i], c_.fork_ev, 0));
  used_ |= 1u << i;
}
Fork::~Fork() {
  // errors here have no caller to reach: a failed join leaves the stream unsynchronised,
  // which the context's next synchronisation reports
  for (int i = 0; i < Ctx::SIDES; ++i)
    if (used_ & (1u << i)) {
      (void)hipEventRecord(c_.side_ev[i], c_.side[i]);
      (void)hipStreamWaitEvent(c_.stream, c_.side_ev[i], 0);
    }
}
void destroy_sides(Ctx& c) {
  for (int i = 0; i < Ctx::SIDES; ++i) {
    if (c.side[i]) {
      (void)hipStreamSynchronize(c.side[i]);
      (void)hipStreamDestroy(c.side[i]);
      (void)hipEventDestroy(c.side_ev[i]);
      c.side[i] = nullptr;
      c.side_ev[i] = nullptr;
    }
  }
  if (c.fork_ev) (void)hipEventDestroy(c.fork_ev);
  c.fork_ev = nullptr;
}

void fill_u32(Ctx& c, uint32_t* p, uint32_t v, uint64_t n) { HGP_LAUNCH(k_fill_u32, ceil_div(n, 256), 256, 0, c, p, v, n); }
void iota_u32(Ctx& c, uint32_t* p, uint64_t n) { HGP_LAUNCH(k_iota_u32, ceil_div(n, 256), 256, 0, c, p, n); }
uint32_t reduce_max_u32(Ctx& c, const uint32_t* p, uint64_t n) {
  DBuf<uint32_t> out(c, 1);
  memset_async(c, out.get(), 0, sizeof(uint32_t));
  if (n) HGP_LAUNCH(k_max_u32, std::min<uint32_t>(ceil_div(n, 256), 2048), 256, 0, c, p, n, out.get());
  return read_scalar(c, out.get());
}

uint64_t reduce_max_u64(Ctx& c, const uint64_t* p, uint64_t n) {
  DBuf<unsigned long long> out(c, 1);
  memset_async(c, out.get(), 0, sizeof(unsigned long long));
  if (n) HGP_LAUNCH(k_max_u64, std::min<uint32_t>(ceil_div(n, 256), 2048), 256, 0, c, p, n, out.get());
  return (uint64_t)read_scalar(c, out.get());
}

}  // namespace hgp
