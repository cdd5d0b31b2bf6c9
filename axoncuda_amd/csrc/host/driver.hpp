// driver.hpp - pipeline orchestration entry points.
#pragma once

#include "../common.hpp"
#include "model.hpp"

namespace hgp {

void initial_partition(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, uint32_t n, uint32_t e,
                       const uint32_t* sizes, uint32_t k, float eps, uint32_t omega, uint32_t*& part_out,
                       uint32_t*& psizes_out, double* ms);

void partition(Ctx& c, const hgp_params& prm, const hgp_hgraph& hg, uint32_t* part_out, hgp_run_stats* stats);

}  // namespace hgp
