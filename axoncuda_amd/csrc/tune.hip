// tune.hip - the autotuner of the library's launch parameters (no counterpart in the
// reference, whose block sizes, hash-set sizes and LDS budgets are compile-time constants
// tuned for one GPU, headers/construction.cuh:22-29, headers/coarsening.cuh:21-31).
//
// Each knob of Ctx::tune changes speed only (every stage returns the same result under any
// valid setting), so the tuner times the candidates of each knob on synthetic inputs built on
// the device and keeps the fastest:
//   * radix-sort tile (keys per workgroup = 256 x rounds) for each size class 2^15 .. 2^max;
//   * the key-count at which scans switch from the single-pass look-back to three passes;
//   * workgroup size of the thread-per-hedge cascade (64 .. 1024);
//   * size and fill limit (occupancy) of the heavy candidate nodes' first-attempt HBM hash
//     tables, the role the reference's -om multiplier plays for its spill hash sets.
// Every candidate is run once to warm up and then timed as the best of three.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "prim.hpp"
#include "stages.hpp"

namespace hgp {
namespace {

__device__ __forceinline__ uint64_t tune_mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void k_tune_keys(uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, uint64_t n, uint32_t mask,
                            uint64_t seed) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = (uint32_t)tune_mix(seed ^ i) & mask;
  if (vals) vals[i] = (uint32_t)i;
}

// hedge e of a synthetic level: 2 + (mix(e) % 15) pins (the thread-per-hedge cascade's range)
__global__ void k_tune_small_len(uint64_t* __restrict__ len, uint32_t e) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < e) len[x] = 2 + tune_mix(x) % 15;
}
__global__ void k_tune_pins(const uint64_t* __restrict__ off, uint32_t e, uint32_t n, uint64_t seed,
                            uint32_t* __restrict__ pins) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= e) return;
  // a local window around a random first pin, like a netlist's nets
  const uint32_t first = (uint32_t)(tune_mix(seed ^ x) % n);
  for (uint64_t i = off[x]; i < off[x + 1]; ++i)
    pins[i] = (uint32_t)((first + tune_mix(seed + i) % 4096) % n);
}
__global__ void k_tune_pm(uint32_t n, uint2* __restrict__ pm2) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint64_t h = tune_mix(v);
  const uint32_t p = (uint32_t)(h & 63), q = (uint32_t)((h >> 8) & 63);
  const bool moves = ((h >> 16) % 10) < 4 && q != p;
  pm2[v] = make_uint2(p | ((moves ? q : 0xFFFFu) << 16), v);  // rank = node id (a permutation)
}
// a heavy level: hub hedges of `len` pins drawn from all n nodes; every hedge also holds
// hub node x % hubs, so the hubs see far more distinct neighbours than LDS tables hold
__global__ void k_tune_hub_pins(uint32_t e, uint32_t len, uint32_t n, uint32_t hubs, uint32_t* __restrict__ pins) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= e) return;
  const uint64_t s = (uint64_t)x * len;
  pins[s] = x % hubs;
  for (uint32_t i = 1; i < len; ++i) pins[s + i] = hubs + (uint32_t)(tune_mix(s + i) % (n - hubs));
}
__global__ void k_tune_offsets(uint64_t* __restrict__ off, uint32_t e, uint32_t len) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x <= e) off[x] = (uint64_t)x * len;
}

struct Timer {
  Ctx& c;
  hipEvent_t a, b;
  explicit Timer(Ctx& cc) : c(cc) {
    HGP_HIP(hipEventCreate(&a));
    HGP_HIP(hipEventCreate(&b));
  }
  ~Timer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  // best of `reps` timed runs of f after one warm-up (which also settles the allocator's
  // size classes); prep runs before each, outside the timed region
  template <typename P, typename F>
  float best_of(P&& prep, F&& f, int reps = 3) {
    prep();
    f();
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
      prep();
      HGP_HIP(hipEventRecord(a, c.stream));
      f();
      HGP_HIP(hipEventRecord(b, c.stream));
      HGP_HIP(hipEventSynchronize(b));
      float ms = 0;
      HGP_HIP(hipEventElapsedTime(&ms, a, b));
      best = std::min(best, ms);
    }
    return best;
  }
};

void tune_sorts(Ctx& c, uint32_t max_log2) {
  Timer t(c);
  const uint64_t nmax = 1ull << max_log2;
  DBuf<uint32_t> k0(c, nmax), v0(c, nmax), k(c, nmax), v(c, nmax);
  HGP_LAUNCH(k_tune_keys, ceil_div(nmax, 256), 256, 0, c, k0.get(), v0.get(), nmax, 0xFFFFFFu, 12345ull);
  for (uint32_t cls = 15; cls <= max_log2; ++cls) {
    const uint64_t n = 1ull << cls;
    uint32_t best_r = 0;
    float best = 1e30f;
    for (uint32_t r = 1; r <= 16; r *= 2) {
      c.tune.sort_rounds[cls] = r;
      const float ms = t.best_of(
          [&] {
            copy_d2d(c, k.get(), k0.get(), n * 4);
            copy_d2d(c, v.get(), v0.get(), n * 4);
          },
          [&] { radix_sort(c, k.get(), v.get(), n, 24); });
      if (ms < best) best = ms, best_r = r;
    }
    c.tune.sort_rounds[cls] = best_r;
    if (c.verbose >= 2) fprintf(stderr, "  tune: sort of 2^%u keys: %u rounds per tile (%.3f ms)\n", cls, best_r, best);
  }
}

void tune_scan(Ctx& c, uint32_t max_log2) {
  Timer t(c);
  constexpr uint64_t TILE = 4096;  // the look-back scan's tile
  uint32_t threshold = 16;
  const uint64_t nmax = std::min<uint64_t>(1ull << max_log2, TILE * 4096);
  DBuf<uint32_t> in(c, nmax);
  DBuf<uint64_t> out(c, nmax + 1);
  HGP_LAUNCH(k_tune_keys, ceil_div(nmax, 256), 256, 0, c, in.get(), (uint32_t*)nullptr, nmax, 0xFFu, 7ull);
  for (uint64_t tiles = 16; tiles * TILE <= nmax; tiles *= 2) {
    const uint64_t n = tiles * TILE;
    c.tune.scan_lookback_tiles = 0xFFFFFFFFu;  // look-back
    const float lb = t.best_of([] {}, [&] { exclusive_scan(c, in.get(), out.get(), n); });
    c.tune.scan_lookback_tiles = 0;  // three passes
    const float tp = t.best_of([] {}, [&] { exclusive_scan(c, in.get(), out.get(), n); });
    if (c.verbose >= 2) fprintf(stderr, "  tune: scan of %llu tiles: look-back %.3f ms, three passes %.3f ms\n",
                                (unsigned long long)tiles, lb, tp);
    if (lb <= tp) threshold = (uint32_t)tiles;
    else break;
  }
  c.tune.scan_lookback_tiles = threshold;
}

void tune_cascade(Ctx& c, uint32_t max_log2) {
  Timer t(c);
  const uint32_t e = 1u << std::min<uint32_t>(max_log2, 22), n = e;
  DBuf<uint64_t> off(c, (uint64_t)e + 1);
  HGP_LAUNCH(k_tune_small_len, ceil_div(e, 256), 256, 0, c, off.get(), e);
  exclusive_scan(c, off.get(), off.get(), e);
  const uint64_t p = read_scalar(c, off.get() + e);
  DBuf<uint32_t> pins(c, p);
  HGP_LAUNCH(k_tune_pins, ceil_div(e, 256), 256, 0, c, off.get(), e, n, 99ull, pins.get());
  DBuf<float> w(c, e);
  DBuf<uint32_t> ones(c, e);
  fill_u32(c, ones.get(), 0x3F800000u, e);  // 1.0f
  copy_d2d(c, w.get(), ones.get(), (uint64_t)e * 4);
  const SmallCsr cs = small_csr(c, nullptr, e, off.get(), pins.get(), w.get());
  DBuf<uint2> pm(c, n);
  HGP_LAUNCH(k_tune_pm, ceil_div(n, 256), 256, 0, c, n, pm.get());
  DBuf<unsigned long long> seq(c, n);
  uint32_t best_b = 256;
  float best = 1e30f;
  for (uint32_t b = 64; b <= 1024; b *= 2) {
    float ms = 1e30f;
    for (int r = 0; r < 4; ++r) ms = std::min(ms, time_cascade_small(c, cs, pm.get(), b, seq.get()));
    if (c.verbose >= 2) fprintf(stderr, "  tune: cascade with %u threads per workgroup: %.3f ms\n", b, ms);
    if (ms < best) best = ms, best_b = b;
  }
  c.tune.cascade_block = best_b;
}

void tune_heavy(Ctx& c) {
  Timer t(c);
  // 2^19 nodes; 8 hubs, each in 256 hedges of 512 pins: about 115 K distinct neighbours per
  // hub (beyond every LDS tier), 2 hedges per other node (the light tier)
  const uint32_t n = 1u << 19;
  const uint32_t hubs = 8, len = 512, e = hubs * 256;
  DBuf<uint64_t> off(c, (uint64_t)e + 1);
  HGP_LAUNCH(k_tune_offsets, ceil_div(e + 1, 256), 256, 0, c, off.get(), e, len);
  DBuf<uint32_t> pins(c, (uint64_t)e * len), srcc(c, e), sizes(c, n);
  HGP_LAUNCH(k_tune_hub_pins, ceil_div(e, 256), 256, 0, c, e, len, n, hubs, pins.get());
  fill_u32(c, srcc.get(), 0u, e);
  fill_u32(c, sizes.get(), 1u, n);
  DBuf<float> w(c, e);
  {
    DBuf<uint32_t> ones(c, e);
    fill_u32(c, ones.get(), 0x3F800000u, e);
    copy_d2d(c, w.get(), ones.get(), (uint64_t)e * 4);
  }
  uint64_t ts;
  uint32_t *tch, *inb;
  uint64_t* toff;
  build_touching(c, pins.get(), off.get(), srcc.get(), n, e, ts, tch, toff, inb);
  DBuf<uint32_t> dt, di;
  DBuf<uint64_t> dto;
  dt.ctx = &c, dt.ptr = tch, dt.n = ts;
  di.ctx = &c, di.ptr = inb, di.n = n;
  dto.ctx = &c, dto.ptr = toff, dto.n = (uint64_t)n + 1;
  const uint32_t cnc = 4;
  DBuf<uint32_t> pairs(c, (uint64_t)n * cnc), scores(c, (uint64_t)n * cnc);
  uint32_t bs = 65536, bf = 12;
  float best = 1e30f;
  for (uint32_t slots = 32768; slots <= 262144; slots *= 2) {
    for (uint32_t f16 = 10; f16 <= 14; f16 += 2) {
      c.tune.heavy_slots = slots;
      c.tune.heavy_fill16 = f16;
      const float ms = t.best_of(
          [] {},
          [&] {
            candidates(c, pins.get(), off.get(), srcc.get(), tch, toff, inb, w.get(), sizes.get(), n, e, cnc,
                       0x7FFFFFFFu, 0x7FFFFFFFu, pairs.get(), scores.get());
          },
          2);
      if (c.verbose >= 2) fprintf(stderr, "  tune: heavy tables of %u slots filled to %u/16: %.3f ms\n", slots, f16, ms);
      if (ms < best) best = ms, bs = slots, bf = f16;
    }
  }
  c.tune.heavy_slots = bs;
  c.tune.heavy_fill16 = bf;
}

}  // namespace

void autotune(Ctx& c, uint32_t max_log2) {
  max_log2 = std::min<uint32_t>(std::max<uint32_t>(max_log2 ? max_log2 : 26, 16), 28);
  tune_sorts(c, max_log2);
  tune_scan(c, max_log2);
  tune_cascade(c, max_log2);
  tune_heavy(c);
  c.tune.tuned = true;
  sync(c);
}

}  // namespace hgp
