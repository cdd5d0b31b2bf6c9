// driver.hpp - pipeline orchestration entry points.
#pragma once

#include "../common.hpp"
#include "../stages.hpp"
#include "model.hpp"

namespace hgp {

// k-way initial partition of a coarse level: Mt-KaHyPar when installed, else the host
// recursive bisection, or the device partitioner when `device_level` (the same level with
// its touching lists) is given.  prune: the keep-target multiplier of the hedge pruning,
// or -1 for the reference's pruning before Mt-KaHyPar and none before the built-in ones.
void initial_partition(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, uint32_t n, uint32_t e,
                       const uint32_t* sizes, uint32_t k, float eps, uint32_t omega, int32_t prune, uint32_t*& part_out,
                       uint32_t*& psizes_out, double* ms, const RefineLevel* device_level = nullptr,
                       const DeviceIpOptions& dip = DeviceIpOptions());

// Level-0 input: either a host model or device-resident arrays (copied into the level).
struct PipelineInput {
  uint32_t n = 0, e = 0;
  uint64_t p = 0;
  const hgp_hgraph* host = nullptr;
  const uint32_t* d_hedges = nullptr;
  const uint64_t* d_hoff = nullptr;
  const uint32_t* d_srcc = nullptr;
  const float* d_w = nullptr;
};

void partition(Ctx& c, const hgp_params& prm, const PipelineInput& in, uint32_t* part_out, hgp_run_stats* stats);

}  // namespace hgp
