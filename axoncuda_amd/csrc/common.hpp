// common.hpp - runtime plumbing shared by every HIP translation unit of libhgp.
//
// * Errors: device/runtime failures throw hgp::Error(status, msg); the C ABI layer
//   (capi.cpp) catches them and returns the status code, so no path exits the process.
// * Memory: stream-ordered allocations (hipMallocAsync on the context stream, pool kept
//   warm) with an HBM budget that every allocation is charged against.
// * Wave64 helpers for gfx950 (CDNA4): lane ids, ballots, shuffles, reductions and scans.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "hgp.h"

namespace hgp {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t WAVE = 64;
constexpr uint32_t FIXED_POINT_SCALE = 262144u;  // 2^18 (reference headers/defines.cuh:24)

struct Error : std::runtime_error {
  hgp_status status;
  Error(hgp_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define HGP_HIP(expr)                                                                             \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess)                                                                         \
      throw ::hgp::Error(HGP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e) + " at " + \
                                        __FILE__ + ":" + std::to_string(__LINE__));               \
  } while (0)

#define HGP_REQUIRE(cond, status, msg)                      \
  do {                                                      \
    if (!(cond)) throw ::hgp::Error((status), (msg));       \
  } while (0)

// One timed launch of the profiled kernel: events around it on the context stream and the
// kernel's algorithmic bytes for that launch.
struct ProfRec {
  hipEvent_t a = nullptr, b = nullptr;
  uint64_t bytes = 0;
};

// Launch parameters chosen by the autotuner (hgp_tuning in the C ABI).  They change speed
// only: every stage returns the same result under any valid setting.
struct Tuning {
  uint32_t sort_rounds[40] = {};     // radix-sort keys per thread and tile, by floor(log2 n); 0 = rule
  uint32_t heavy_slots = 65536;      // first-attempt HBM hash table of a heavy candidate node
  uint32_t heavy_fill16 = 12;        // its fill limit in sixteenths
  uint32_t cascade_block = 256;      // threads per workgroup of the thread-per-hedge cascade
  uint32_t scan_lookback_tiles = 64;  // largest scan (in tiles) that takes the single-pass look-back
  uint32_t mask_wide4_degree = 128;   // mean touching hedges per node from which k-way gains take
  uint32_t mask_wide16_degree = 512;  // 4 (16) waves per node
  uint32_t side_streams = 2;          // bit 0: candidate tiers, bit 1: cascade classes beside events,
                                      // bit 2: ... and beside the small-hedge cascade
  uint32_t coop_max_n = 65536;        // sorts of RS1 .. this many keys run every pass in one launch
  uint32_t coop_tile = 1024;          // keys per workgroup of those launches (at least)
  bool tuned = false;
};

// Execution context: device, stream, HBM budget accounting, last error text.
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  uint64_t budget = 0;  // 0 = whole device
  uint64_t in_use = 0;   // bytes handed out (size classes)
  uint64_t peak = 0;
  uint64_t cached = 0;   // bytes parked in the free lists
  std::unordered_map<void*, uint64_t> live;                    // pointer -> size class
  std::unordered_map<uint64_t, std::vector<void*>> free_lists;  // size class -> parked blocks
  std::string last_error;
  int verbose = 0;
  uint32_t num_cus = 256;
  // status arrays of the single-pass scans (prim.hip)
  uint32_t* scan_flags = nullptr;
  void* scan_agg = nullptr;
  void* scan_incl = nullptr;
  unsigned long long* scan_ticket = nullptr;
  unsigned int* coop_bar = nullptr;  // grid-barrier counter of the cooperative sorts (prim.hip)
  uint64_t coop_issued = 0;
  uint32_t scan_cap_tiles = 0;
  uint32_t scan_gen = 0;
  uint64_t scan_tickets_issued = 0;
  std::string prof_kernel;     // base name of the kernel whose launches are timed ("" = none)
  std::vector<ProfRec> prof;   // timed launches not yet read
  // algorithmic bytes that the profiled kernel counts itself on the device (PROF_SLOTS
  // counters, spread so that the atomics do not all meet on one address)
  unsigned long long* prof_dev = nullptr;
  bool prof_count = false;  // kernels with a byte model count their bytes (extra device work)
  void* pinned = nullptr;  // small pinned host buffer for scalar read-backs (allocated on first use)
  Tuning tune;
  // device bytes the refinement's inbound-set event stage may take at once (0: half of what
  // the device has free, cached blocks included)
  uint64_t event_budget = 0;
  // side streams for independent kernels of one stage (created on first use, default priority)
  static constexpr int SIDES = 3;
  hipStream_t side[SIDES] = {};
  hipEvent_t fork_ev = nullptr, side_ev[SIDES] = {};
};

// Independent kernels of one stage on the context's side streams, beside the main stream's.
// The side streams start after everything queued on the main stream so far; when the Fork
// ends, the main stream waits for all work given to them.  Work given to a side stream must
// not allocate or free device memory (the caching allocator is ordered by the main stream) nor
// use the context's scans or sorts (their status arrays are shared), and the buffers it reads
// or writes must outlive the Fork.
class Fork {
 public:
  explicit Fork(Ctx& c, bool enabled = true);
  ~Fork();
  template <typename F>
  void on(int i, F&& f) {  // f's launches go to side stream i (the main stream when disabled)
    if (!enabled_) {
      f();
      return;
    }
    start(i);
    struct Restore {
      Ctx& c;
      hipStream_t s;
      ~Restore() { c.stream = s; }
    } restore{c_, c_.stream};
    c_.stream = c_.side[i];
    f();
  }
  Fork(const Fork&) = delete;
  Fork& operator=(const Fork&) = delete;

 private:
  void start(int i);
  Ctx& c_;
  bool enabled_;
  uint32_t used_ = 0;
};
void destroy_sides(Ctx& c);
constexpr uint64_t PINNED_BYTES = 256;
constexpr uint32_t PROF_SLOTS = 64;

// Device memory: a caching allocator on top of hipMallocAsync.  Freed blocks are parked
// per size class (4 classes per power of two) and handed out again by later requests on
// the same stream, which is safe in stream order and avoids a HIP call per scratch buffer.
void* dmalloc(Ctx& c, uint64_t bytes);
void dfree(Ctx& c, void* p);
void release_cached(Ctx& c);  // returns every parked block to HIP

// RAII device buffer tied to a context.
template <typename T>
struct DBuf {
  Ctx* ctx = nullptr;
  T* ptr = nullptr;
  uint64_t n = 0;
  DBuf() = default;
  DBuf(Ctx& c, uint64_t count) : ctx(&c), n(count) {
    ptr = count ? (T*)dmalloc(c, count * sizeof(T)) : nullptr;
  }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : ctx(o.ctx), ptr(o.ptr), n(o.n) {
    o.ptr = nullptr;
    o.n = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    reset();
    ctx = o.ctx;
    ptr = o.ptr;
    n = o.n;
    o.ptr = nullptr;
    o.n = 0;
    return *this;
  }
  ~DBuf() { reset(); }
  void reset() {
    if (ptr && ctx) dfree(*ctx, ptr);
    ptr = nullptr;
    n = 0;
  }
  T* release() {
    T* p = ptr;
    ptr = nullptr;
    n = 0;
    return p;
  }
  T* get() const { return ptr; }
  operator T*() const { return ptr; }
};

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

inline void memset_async(Ctx& c, void* p, int v, uint64_t bytes) {
  if (bytes) HGP_HIP(hipMemsetAsync(p, v, bytes, c.stream));
}
inline void copy_d2d(Ctx& c, void* dst, const void* src, uint64_t bytes) {
  if (bytes) HGP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c.stream));
}
inline void copy_h2d(Ctx& c, void* dst, const void* src, uint64_t bytes) {
  if (bytes) HGP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream));
}
// Small read-backs land in the context's pinned buffer (a direct copy, no staging through
// pageable memory), larger ones straight in `dst`; either way the stream is synchronized.
inline void copy_d2h(Ctx& c, void* dst, const void* src, uint64_t bytes) {
  if (!bytes) return;
  if (bytes <= PINNED_BYTES) {
    if (!c.pinned) HGP_HIP(hipHostMalloc(&c.pinned, PINNED_BYTES, hipHostMallocDefault));
    HGP_HIP(hipMemcpyAsync(c.pinned, src, bytes, hipMemcpyDeviceToHost, c.stream));
    HGP_HIP(hipStreamSynchronize(c.stream));
    std::memcpy(dst, c.pinned, bytes);
    return;
  }
  HGP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
  HGP_HIP(hipStreamSynchronize(c.stream));
}
template <typename T>
inline T read_scalar(Ctx& c, const T* d) {
  T v;
  copy_d2h(c, &v, d, sizeof(T));
  return v;
}
inline void sync(Ctx& c) { HGP_HIP(hipStreamSynchronize(c.stream)); }

// True when `spelled` (the launch site's kernel expression, e.g. "(k_rs_scatter<K, true>)")
// names the profiled kernel: its identifier before any template arguments.
inline bool prof_match(const Ctx& c, const char* spelled) {
  if (c.prof_kernel.empty()) return false;
  while (*spelled == '(' || *spelled == ' ') ++spelled;
  size_t len = 0;
  while (spelled[len] && spelled[len] != '<' && spelled[len] != ')' && spelled[len] != ' ') ++len;
  return c.prof_kernel.size() == len && c.prof_kernel.compare(0, len, spelled, len) == 0;
}
inline hipEvent_t prof_start(Ctx& c) {
  hipEvent_t e;
  HGP_HIP(hipEventCreate(&e));
  HGP_HIP(hipEventRecord(e, c.stream));
  return e;
}
inline void prof_stop(Ctx& c, hipEvent_t a, uint64_t bytes) {
  hipEvent_t e;
  HGP_HIP(hipEventCreate(&e));
  HGP_HIP(hipEventRecord(e, c.stream));
  c.prof.push_back({a, e, bytes});
}

// Kernel launch on the context stream.  HGP_LAUNCH_B also states the launch's algorithmic
// bytes, which the kernel profiler (hgp_profile_kernel) reports next to the timing.
#define HGP_LAUNCH_B(bytes, kernel, grid, block, shmem, ctx, ...)                             \
  do {                                                                                      \
    if ((grid) > 0) {                                                                       \
      const bool _prof = ::hgp::prof_match((ctx), #kernel);                                 \
      const uint64_t _bytes = _prof ? (uint64_t)(bytes) : 0; /* before the timed launch */  \
      hipEvent_t _pa = _prof ? ::hgp::prof_start(ctx) : nullptr;                            \
      hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, (ctx).stream, __VA_ARGS__); \
      HGP_HIP(hipGetLastError());                                                           \
      if (_prof) ::hgp::prof_stop((ctx), _pa, _bytes);                                      \
    }                                                                                       \
  } while (0)
#define HGP_LAUNCH(kernel, grid, block, shmem, ctx, ...) HGP_LAUNCH_B(0, kernel, grid, block, shmem, ctx, __VA_ARGS__)

// ---------------------------------------------------------------------------------------
// device helpers (wave64)
#if defined(__HIPCC__)

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// number of set bits of `mask` strictly below this lane
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Wave-wide reductions (all 64 lanes active; every lane gets the result).  The first four
// exchanges are DPP lane moves inside 16-lane rows (quad xor 1, quad xor 2, half-row mirror,
// row mirror: each pairs disjoint groups, so sums stay exact), the fifth a ds_swizzle that
// swaps the 16-lane halves of each 32, and the two 32-lane results are read as scalars.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_move(T v) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = (uint64_t)v;
    return (T)(((uint64_t)dpp_u32<CTRL>((uint32_t)(u >> 32)) << 32) | dpp_u32<CTRL>((uint32_t)u));
  } else {
    return (T)dpp_u32<CTRL>((uint32_t)v);
  }
}
template <typename T>
__device__ __forceinline__ T swap16_move(T v) {  // lane i <- lane i ^ 16
  auto sw = [](uint32_t x) { return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x401F); };
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = (uint64_t)v;
    return (T)(((uint64_t)sw((uint32_t)(u >> 32)) << 32) | sw((uint32_t)u));
  } else {
    return (T)sw((uint32_t)v);
  }
}
template <typename T>
__device__ __forceinline__ T lane_value(T v, int lane) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return (T)(((uint64_t)hi << 32) | lo);
  } else {
    return (T)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  }
}
template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
  v = op(v, dpp_move<0xB1>(v));   // quad_perm [1,0,3,2]: lane ^ 1
  v = op(v, dpp_move<0x4E>(v));   // quad_perm [2,3,0,1]: lane ^ 2
  v = op(v, dpp_move<0x141>(v));  // row_half_mirror: the two quads of each 8
  v = op(v, dpp_move<0x140>(v));  // row_mirror: the two halves of each 16
  v = op(v, swap16_move(v));      // the two rows of each 32
  return op(lane_value(v, 0), lane_value(v, 32));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  return wave_reduce(v, [](T a, T b) { return (T)(a + b); });
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  return wave_reduce(v, [](T a, T b) { return b > a ? b : a; });
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
  return wave_reduce(v, [](T a, T b) { return b < a ? b : a; });
}

// Workgroup b of a grid of g is dispatched to XCD b % 8 (round-robin over the 8 XCDs).  The
// returned logical index hands XCD x one contiguous range of [0, g), so work items of nearby
// ids (nodes that share hedges) meet in the same XCD's L2 instead of being fetched into all
// eight.  A bijection on [0, g): XCD x runs g / 8 workgroups, one more when x < g % 8.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t g) {
  constexpr uint32_t XCDS = 8;
  const uint32_t x = b % XCDS, i = b / XCDS, q = g / XCDS, r = g % XCDS;
  return x * q + (x < r ? x : r) + i;
}

// inclusive prefix sum across the 64 lanes
template <typename T>
__device__ __forceinline__ T wave_scan_incl(T v) {
  const uint32_t l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T x = __shfl_up(v, o, 64);
    if (l >= (uint32_t)o) v += x;
  }
  return v;
}

// 2^18 fixed point of a hedge weight, saturated to [0, 2^32 - 1]
__device__ __forceinline__ uint32_t fixed_weight(float w) {
  float v = w * 262144.0f;
  if (!(v > 0.0f)) return 0u;
  if (v >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)v;
}

// Symmetric pairing noise in [0, 64): the mixing steps of the reference's
// deterministic_noise<64> (headers/utils.cuh:159-171).
__device__ __forceinline__ uint32_t noise64(uint32_t a, uint32_t b) {
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  uint32_t x = lo * 0x9E3779B1u;
  x ^= hi + 0x85EBCA6Bu + (x << 6) + (x >> 2);
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 13;
  x *= 0x9E3779B1u;
  x ^= x >> 16;
  return x & 63u;
}

// Integer mixer used for gain tie-breaks: the steps of the reference's hash_uint32
// (headers/utils.cuh:252-261).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 17;
  x *= 0xED5AD4BBu;
  x ^= x >> 11;
  x *= 0xAC4C1B51u;
  x ^= x >> 15;
  x *= 0x31848BABu;
  x ^= x >> 14;
  return x;
}

__device__ __forceinline__ uint32_t slot_hash(uint32_t x) { return x * 0x9E3779B1u; }

template <typename T>
__device__ __forceinline__ T atomic_load_relaxed(T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

#endif

}  // namespace hgp
