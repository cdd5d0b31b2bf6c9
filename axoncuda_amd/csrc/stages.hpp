// stages.hpp - host-side entry points of the pipeline stages (device pointers in/out).
// Outputs returned through `T*&` are allocated with dmalloc on the context and owned by
// the caller (released with dfree), matching the reference's step-function contract.
#pragma once

#include "common.hpp"

namespace hgp_ip {
struct Hypergraph;
}

namespace hgp {

// ---- construction (construct.hip) -----------------------------------------------------
void build_touching(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const uint32_t* srcc, uint32_t n,
                    uint32_t e, uint64_t& tsize, uint32_t*& tch, uint64_t*& toff, uint32_t*& inb);
void build_neighbors(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const uint32_t* tch, const uint64_t* toff,
                     uint32_t n, uint64_t& nsize, uint32_t*& nbr, uint64_t*& noff);
void coarsen_neighbors(Ctx& c, const uint32_t* nbr, const uint64_t* noff, const uint32_t* groups, uint32_t n,
                       uint32_t new_n, uint64_t& nsize, uint32_t*& cnbr, uint64_t*& cnoff);
void coarsen_hedges(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const uint32_t* srcc, const uint32_t* groups,
                    uint32_t e, uint64_t hsize, uint32_t new_n, uint64_t& csize, uint32_t*& chedges, uint64_t*& choff,
                    uint32_t*& csrcc);
void uncoarsen(Ctx& c, const uint32_t* groups, const uint32_t* cpart, uint32_t n, uint32_t* part);

// ---- coarsening (coarsen.hip) ---------------------------------------------------------
void candidates(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const uint32_t* srcc, const uint32_t* tch,
                const uint64_t* toff, const uint32_t* inb, const float* w, const uint32_t* sizes, uint32_t n,
                uint32_t e, uint32_t cnc, uint32_t omega, uint32_t delta, uint32_t* pairs, uint32_t* scores);
void group_nodes(Ctx& c, const uint32_t* inb, const uint32_t* pairs, const uint32_t* scores, const uint32_t* sizes,
                 uint32_t n, uint32_t cnc, uint32_t omega, uint32_t delta, uint32_t& new_n, uint32_t*& groups,
                 uint32_t*& gsizes, uint32_t** ungroups, uint64_t** uoff, uint32_t* raw_out = nullptr);

// ---- refinement (refine.hip) ------------------------------------------------------------
struct RefineLevel {
  uint32_t n, e, k;
  const uint32_t* hedges;
  const uint64_t* hoff;
  const uint32_t* srcc;
  const uint32_t* tch;
  const uint64_t* toff;
  const uint32_t* inb;
  const float* w;
  const uint32_t* sizes;
};
struct AnalyzeOut {  // optional host copies of the intermediates
  uint32_t* moves = nullptr;
  int64_t* gains = nullptr;
  uint32_t* ranks = nullptr;
  int64_t* seq = nullptr;
  int32_t* sviol = nullptr;
  int32_t* iviol = nullptr;
};
void refine(Ctx& c, const RefineLevel& L, int mode, uint32_t omega, uint32_t delta, uint32_t repeats,
            bool final_counts, uint32_t* part, uint32_t* psizes, uint32_t* pinbound);
uint32_t refine_analyze(Ctx& c, const RefineLevel& L, int mode, uint32_t omega, uint32_t delta, uint32_t repeats,
                        uint32_t fm_repeat, bool chainup, const uint32_t* part, const uint32_t* psizes,
                        const AnalyzeOut& out);
// Device k-way initial partition of a level (best-fit-decreasing starts, Jacobi rounds, the
// start with the lowest lambda-1 kept): allocates part[n] and psizes[k].
struct DeviceIpOptions {
  uint32_t starts = 2;           // best-fit-decreasing starts
  uint32_t rounds = 8;           // Jacobi rounds per start (fewer when a round applies nothing)
  uint32_t nodes_per_part = 16;  // coarsen the level further to about this many nodes per part
  uint32_t refine_repeats = 4;   // k-way refinement rounds per projected level
};
void initial_partition_device(Ctx& c, const RefineLevel& L, uint32_t omega, const DeviceIpOptions& opt,
                              uint32_t prune, uint32_t*& part_out, uint32_t*& psizes_out);
// One Jacobi label-propagation round (k-way): every node proposes its best move, positive-
// gain moves are sorted by (target, gain desc, node) and applied while they fit next to the
// target's size before the round.  Returns the number of moves applied.
uint32_t jacobi_round(Ctx& c, const RefineLevel& L, uint32_t omega, uint32_t round, uint32_t* part, uint32_t* psizes);
// Label-propagation refinement (k-way): up to `rounds` Jacobi rounds; a round that applies
// nothing ends them, a round that raises lambda-1 is undone and ends them.
void lp_refine(Ctx& c, const RefineLevel& L, uint32_t omega, uint32_t rounds, uint32_t* part, uint32_t* psizes);
// Hedges of 2..16 pins as the thread-per-hedge cascade reads them: each hedge's pins start at a
// multiple of 4 and are padded to one with its last pin; rec = start << 5 | length.
struct SmallCsr {
  DBuf<uint64_t> rec;
  DBuf<uint32_t> pins;
  DBuf<float> w;
  uint32_t m = 0;
};
// The small-hedge CSR of hedges keys[i] (low 32 bits; i itself without keys), i < m, of the CSR
// (hedges, hoff) with weights w.
SmallCsr small_csr(Ctx& c, const uint64_t* keys, uint32_t m, const uint64_t* hoff, const uint32_t* hedges,
                   const float* w);
// Milliseconds of one thread-per-hedge cascade launch over a small-hedge CSR with `block`
// threads per workgroup (pm2: per node partition | target << 16 (0xFFFF: no move), rank; seq:
// one counter per rank).  For the autotuner.
float time_cascade_small(Ctx& c, const SmallCsr& cs, const uint2* pm2, uint32_t block, unsigned long long* seq);
// The autotuner (tune.hip): times the launch parameters of Ctx::tune on synthetic inputs of
// up to 2^max_log2 keys / hedges and keeps the fastest of each in c.tune.
void autotune(Ctx& c, uint32_t max_log2);
void inbound_sizes(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const uint32_t* srcc, const uint32_t* part,
                   uint32_t e, uint32_t k, uint32_t* out);
int64_t connectivity_fixed(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, const uint32_t* part,
                           uint32_t e);

// ---- lambda-1 without device allocation (noalloc.hip) -----------------------------------
constexpr uint32_t NOALLOC_MAX_PARTS = 65536;  // one LDS bitmap of this many bits per wave
constexpr uint32_t NOALLOC_MAX_DEVICES = 64;
// Looks up this device's copy of the file-scope scratch array (the device must be current).
void noalloc_setup(int device);
// *d_out = sum_e W_e (lambda(e) - 1) in 2^18 fixed point, or -1 when a pin's partition id is
// >= k; launches two kernels on the context stream and allocates nothing.
void connectivity_noalloc(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, const uint32_t* part,
                          uint32_t e, uint32_t k, int64_t* d_out);

// ---- initial partitioner input (ipprep.hip) ------------------------------------------------
// The coarsest level as the initial partitioner takes it (hedges of >= 2 distinct pins,
// sorted, identical sets merged in order of first appearance with summed 2^18 weights,
// incidence lists), built on the device.  With `prune` = f > 0, hedges are first dropped or
// kept with rescaled weights by the reference's degree-score pruning (sources/init_part.cu:749-817)
// aiming at f times its target of N * floor(P / E) kept hedges.
void ip_input_device(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, const uint32_t* sizes,
                     uint32_t n, uint32_t e, uint32_t prune, hgp_ip::Hypergraph& h);
// The same merge left on the device as a level (sorted pins, no sources): e_out hedges with
// float weights = summed fixed-point weights / 2^18.
void merge_level_device(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const float* w, uint32_t n, uint32_t e,
                        uint32_t prune, uint32_t& e_out, uint32_t*& hedges_out, uint64_t*& hoff_out, float*& w_out);

// ---- quotient hypergraph (quotient.hip) ---------------------------------------------------
// getPartitionsHypergraph(part, 2, true) on the device: k_out partitions (ids must be 0..k-1,
// all used), e_out merged hedges of p_out pins in order of first appearance (sources then
// destinations, each ascending) with their float weights summed in hedge order.
void quotient(Ctx& c, const uint32_t* hedges, const uint64_t* hoff, const uint32_t* srcc, const float* w,
              const uint32_t* part, uint32_t n, uint32_t e, uint32_t& k_out, uint32_t& e_out, uint64_t& p_out,
              uint32_t*& qpins, uint64_t*& qoff, uint32_t*& qsrcc, float*& qw);

// ---- post-processing (postproc.hip) -----------------------------------------------------
void greedy_merge_groups(Ctx& c, const uint32_t* sizes, const uint32_t* inb, uint32_t n, uint32_t ngroups,
                         uint32_t omega, uint32_t delta, uint32_t* groups, uint32_t* gsizes, uint32_t& new_ngroups);
void merge_small_partitions(Ctx& c, const uint32_t* psizes, const uint32_t* pinbound, uint32_t n, uint32_t k,
                            uint32_t omega, uint32_t delta, uint32_t* part);
void relabel(Ctx& c, uint32_t n, uint32_t* part, uint32_t& new_k);


}  // namespace hgp
