// noalloc.hip - lambda-1 on the device with no device allocation.
//
// The same figure as hgp_connectivity (sum_e W_e * (lambda(e) - 1) in 2^18 fixed point,
// compute_connectivity_kernel, kernels/init_part.cu:287-314), for callers that cannot
// allocate device memory: code under HIP graph capture, tight serving loops.  Instead of
// building per-hedge partition lists, every wave counts the distinct partitions of one hedge
// at a time:
//   * hedges of at most 64 pins: one pin per lane, each lane compares its partition with the
//     lanes before it (64 broadcasts at most), lambda = number of first occurrences;
//   * longer hedges: a bitmap of NOALLOC_MAX_PARTS bits per wave in LDS; each pin sets its
//     partition's bit (an LDS atomic OR, counting the bits it newly set), a second pass over
//     the pins clears them again.
// Partial sums of the workgroups, a ticket counter and an error flag live in the file-scope
// array `scratch` (NOALLOC_SCRATCH_WORDS 64-bit words); the last workgroup to take a ticket
// adds the partials and writes the result.  Each device holds its own copy of the array, at
// the address noalloc_setup() looks up once per device.  Calls that overlap on one device
// share that array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "stages.hpp"

namespace hgp {
namespace {

constexpr uint32_t NOALLOC_BLOCKS = 2048;  // workgroups of the counting kernel at most
constexpr uint32_t NOALLOC_WAVES = 4;      // waves per workgroup
constexpr uint32_t NOALLOC_WORDS = NOALLOC_MAX_PARTS / 32;  // bitmap words per wave
// scratch layout (64-bit words): [0, NOALLOC_BLOCKS) partial sums, then the ticket and the
// out-of-range flag
constexpr uint32_t NOALLOC_TICKET = NOALLOC_BLOCKS;
constexpr uint32_t NOALLOC_FLAG = NOALLOC_BLOCKS + 1;
constexpr uint32_t NOALLOC_SCRATCH_WORDS = NOALLOC_BLOCKS + 64;

static __device__ __align__(256) unsigned long long scratch[NOALLOC_SCRATCH_WORDS];

std::mutex g_setup_mu;
unsigned long long* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// ticket and flag start at zero on every call (a kernel, not a memset: the call may be captured)
__global__ void k_noalloc_clear(unsigned long long* __restrict__ scr) {
  if (threadIdx.x == 0) {
    scr[NOALLOC_TICKET] = 0;
    scr[NOALLOC_FLAG] = 0;
  }
}

__global__ void __launch_bounds__(64 * NOALLOC_WAVES)
    k_noalloc_km1(const uint32_t* __restrict__ hedges, const uint64_t* __restrict__ hoff, const float* __restrict__ w,
                  const uint32_t* __restrict__ part, uint32_t E, uint32_t K, unsigned long long* __restrict__ scr,
                  int64_t* __restrict__ out) {
  __shared__ uint32_t bm[NOALLOC_WAVES][NOALLOC_WORDS];
  __shared__ int64_t wsum[NOALLOC_WAVES];
  __shared__ bool last;
  const uint32_t wv = threadIdx.x / 64, lane = lane_id();
  uint32_t* my = bm[wv];
  for (uint32_t i = lane; i < NOALLOC_WORDS; i += 64) my[i] = 0;
  __builtin_amdgcn_wave_barrier();
  int64_t acc = 0;
  bool bad = false;
  for (uint32_t e = blockIdx.x * NOALLOC_WAVES + wv; e < E; e += gridDim.x * NOALLOC_WAVES) {
    const uint64_t s = hoff[e], t = hoff[e + 1];
    const uint64_t len = t - s;
    uint32_t lambda = 0;
    if (len <= 64) {
      const bool in = lane < len;
      const uint32_t q = in ? part[hedges[s + lane]] : NONE;
      bad |= in && q >= K;
      bool first = in;
      for (uint32_t j = 0; j < (uint32_t)len; ++j) {
        const uint32_t qj = __shfl(q, (int)j, 64);
        first &= !(j < lane && qj == q);
      }
      lambda = (uint32_t)__popcll(__ballot(first));
    } else {
      uint32_t fresh = 0;
      for (uint64_t i = s + lane; i < t; i += 64) {
        const uint32_t q = part[hedges[i]];
        if (q >= K) {
          bad = true;
          continue;
        }
        const uint32_t bit = 1u << (q & 31);
        fresh += (atomicOr(&my[q >> 5], bit) & bit) ? 0u : 1u;
      }
      lambda = (uint32_t)wave_sum((uint64_t)fresh);
      __builtin_amdgcn_wave_barrier();
      for (uint64_t i = s + lane; i < t; i += 64) {
        const uint32_t q = part[hedges[i]];
        if (q < K) my[q >> 5] = 0;
      }
      __builtin_amdgcn_wave_barrier();
    }
    if (lambda > 1) acc += (int64_t)fixed_weight(w[e]) * (int64_t)(lambda - 1);
  }
  if (__ballot(bad) && lane == 0) atomicOr(&scr[NOALLOC_FLAG], 1ull);
  if (lane == 0) wsum[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t b = 0;
    for (uint32_t i = 0; i < NOALLOC_WAVES; ++i) b += wsum[i];
    scr[blockIdx.x] = (unsigned long long)b;
    __threadfence();
    last = atomicAdd(&scr[NOALLOC_TICKET], 1ull) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  int64_t tot = 0;
  for (uint32_t i = threadIdx.x; i < gridDim.x; i += blockDim.x)
    tot += (int64_t)__hip_atomic_load(&scr[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  tot = (int64_t)wave_sum((uint64_t)tot);
  if (lane == 0) wsum[wv] = tot;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t r = 0;
    for (uint32_t i = 0; i < NOALLOC_WAVES; ++i) r += wsum[i];
    const bool flagged = __hip_atomic_load(&scr[NOALLOC_FLAG], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    *out = flagged ? -1 : r;
  }
}

}  // namespace

void noalloc_setup(int device) {
  HGP_REQUIRE(device >= 0 && device < (int)NOALLOC_MAX_DEVICES, HGP_E_CAPACITY,
              "no-allocation scratch: device index beyond the supported count");
  std::lock_guard<std::mutex> lk(g_setup_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  // the caller has made `device` current; every device holds its own copy of `scratch`
  HGP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  HGP_REQUIRE(p != nullptr, HGP_E_HIP, "no-allocation scratch: symbol address not found");
  g_scratch_ptr[device] = (unsigned long long*)p;
}

void connectivity_noalloc(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, const uint32_t* part,
                          uint32_t e, uint32_t k, int64_t* d_out) {
  HGP_REQUIRE(k <= NOALLOC_MAX_PARTS, HGP_E_CAPACITY,
              "hgp_connectivity_noalloc: num_partitions exceeds the limit of 65536 (one LDS bitmap per wave)");
  unsigned long long* scr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_setup_mu);
    if (c.device >= 0 && c.device < (int)NOALLOC_MAX_DEVICES) scr = g_scratch_ptr[c.device];
  }
  HGP_REQUIRE(scr != nullptr, HGP_E_INTERNAL, "hgp_connectivity_noalloc: scratch not set up for this device");
  const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(e, NOALLOC_WAVES), NOALLOC_BLOCKS));
  HGP_LAUNCH(k_noalloc_clear, 1, 64, 0, c, scr);
  HGP_LAUNCH(k_noalloc_km1, blocks, 64 * NOALLOC_WAVES, 0, c, hedges, hoff, w, part, e, k, scr, d_out);
}

}  // namespace hgp
