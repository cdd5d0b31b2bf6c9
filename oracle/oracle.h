/*
 * oracle.h - CPU restatement of the reference hgraph_gpu pipeline (test infrastructure).
 *
 * This library is the parity checker for the HIP implementation.  It is used only by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg; the product
 * never links or calls it.  Each function is a plain sequential transcription of one
 * reference step (file:line cited at the definition) restated on host arrays, in the
 * canonical output orders listed in include/hgp.h.
 *
 * Conventions: arrays are caller-allocated host buffers; functions return 0 on success.
 * Two-phase functions accept a NULL output array to only compute sizes/offsets.
 */
#ifndef HGP_ORACLE_H
#define HGP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Loads .hgr/.snn/.axh (by extension) and applies rule-2 dedup when `dedup` != 0.
 * Returns a handle; query with orc_hg_sizes / orc_hg_copy, release with orc_hg_free. */
void* orc_hg_load(const char* path, int dedup);
void orc_hg_sizes(const void* h, uint32_t* n, uint32_t* e, uint64_t* p);
void orc_hg_copy(const void* h, uint64_t* hoff, uint32_t* pins, uint32_t* srcc, float* w);
void orc_hg_free(void* h);

/* deduplicateHyperedges(2,false): in-place on a CSR; returns the new pin count. */
uint64_t orc_dedup(uint32_t e, uint64_t* hoff, uint32_t* pins, uint32_t* srcc);

/* 2^18 fixed point of a hedge weight, saturated to [0, 2^32-1]. */
uint32_t orc_fixed_weight(float w);
uint32_t orc_noise(uint32_t a, uint32_t b);
uint32_t orc_hash_uint32(uint32_t x);

int orc_build_touching(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins,
                       const uint32_t* srcc, uint64_t* toff, uint32_t* tch, uint32_t* inb);

int orc_build_neighbors(uint32_t n, const uint64_t* hoff, const uint32_t* pins, const uint64_t* toff,
                        const uint32_t* tch, uint64_t* noff, uint32_t* nbr);

int orc_candidates(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                   const uint64_t* toff, const uint32_t* tch, const uint32_t* inb, const float* w,
                   const uint32_t* sizes, uint32_t cnc, uint32_t omega, uint32_t delta, uint32_t* pairs,
                   uint32_t* scores);

/* Pairing + orphans + relabel.  groups[n]; group_sizes[n] (first *new_n valid);
 * ungroups[n]; uoff[n+1] (first *new_n+1 valid).  raw_groups[n] (nullable) receives the
 * pre-relabel group ids. */
int orc_group_nodes(uint32_t n, const uint32_t* inb, const uint32_t* pairs, const uint32_t* scores,
                    const uint32_t* sizes, uint32_t cnc, uint32_t omega, uint32_t delta, uint32_t* new_n,
                    uint32_t* groups, uint32_t* group_sizes, uint32_t* ungroups, uint64_t* uoff,
                    uint32_t* raw_groups);

int orc_coarsen_hedges(uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                       const uint32_t* groups, uint64_t* choff, uint32_t* cpins, uint32_t* csrcc);

int orc_coarsen_neighbors(uint32_t n, const uint64_t* noff, const uint32_t* nbr, const uint32_t* groups,
                          uint32_t new_n, uint64_t* cnoff, uint32_t* cnbr);

/* One refinement analysis round (see hgp_refine_analyze). */
int orc_refine_analyze(int mode, uint32_t omega, uint32_t delta, uint32_t refine_repeats, uint32_t n,
                       uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                       const uint64_t* toff, const uint32_t* tch, const uint32_t* inb, const float* w,
                       const uint32_t* sizes, uint32_t k, uint32_t fm_repeat, int chainup,
                       const uint32_t* part, const uint32_t* psizes, uint32_t* moves, int64_t* gains,
                       uint32_t* ranks, int64_t* seq_gains, int32_t* size_viol, int32_t* inbound_viol,
                       uint32_t* best_rank);

/* refinementRepeats: updates part/psizes; writes pinbound (k) when final != 0. */
int orc_refine(int mode, uint32_t omega, uint32_t delta, uint32_t refine_repeats, uint32_t n, uint32_t e,
               const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc, const uint64_t* toff,
               const uint32_t* tch, const uint32_t* inb, const float* w, const uint32_t* sizes, uint32_t k,
               int final_counts, uint32_t* part, uint32_t* psizes, uint32_t* pinbound);

int orc_inbound_sizes(uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                      const uint32_t* part, uint32_t k, uint32_t* out);

int64_t orc_connectivity_fixed(uint32_t e, const uint64_t* hoff, const uint32_t* pins, const float* w,
                               const uint32_t* part);

int orc_greedy_merge_groups(uint32_t n, uint32_t ngroups, const uint32_t* sizes, const uint32_t* inb,
                            uint32_t omega, uint32_t delta, uint32_t* groups, uint32_t* group_sizes,
                            uint32_t* new_ngroups);

int orc_merge_small_partitions(uint32_t n, uint32_t k, const uint32_t* psizes, const uint32_t* pinbound,
                               uint32_t omega, uint32_t delta, uint32_t* part);

int orc_relabel(uint32_t n, uint32_t* part, uint32_t* new_k);

/* 1: the built-in initial partitioner runs every try and half on the calling thread;
 * 0 (default): it uses up to 16 threads.  The partition is the same either way. */
void orc_set_ip_threads(int threads);

/* Chaining alone (sources/chaining.cu:19-254): ranks plus the links behind them, per node:
 * predecessor (UINT32_MAX for a head), component head and position along the chain. */
int orc_chain_links(uint32_t n, const uint32_t* part, const uint32_t* moves, const int64_t* gains,
                    const uint32_t* sizes, uint32_t* ranks, uint32_t* prev, uint32_t* comp, uint32_t* pos);

/* The initial partitioner's input (sources/init_part.cu:749-888), restated here: optional
 * hedge pruning, then hedges of >= 2 pins, sorted, identical sets merged in order of first
 * appearance with summed weights; two-phase (NULL arrays return the sizes). */
int orc_ip_input(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const float* w,
                 const uint32_t* sizes, int prune, uint32_t* e_out, uint64_t* p_out, uint64_t* off_out,
                 uint32_t* pins_out, uint64_t* w_out, uint64_t* vw_out);

/* Built-in k-way initial partitioner on a level: the oracle's own input restatement (with
 * the hedge pruning when `prune` = f > 0, aiming at f * N * floor(P / E) hedges), partitioned by the product's recursive bisection. */
int orc_initial_partition(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins,
                          const float* w, const uint32_t* sizes, uint32_t k, float eps, uint32_t omega,
                          int prune, uint32_t* part, uint32_t* psizes);

/* One Jacobi label-propagation round, k-way, on a level whose hedges are all destinations
 * (the device initial partitioner's step and the optional LP refinement); *applied = moves. */
int orc_jacobi_round(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint64_t* toff,
                     const uint32_t* tch, const uint32_t* inb, const float* w, const uint32_t* sizes, uint32_t k,
                     uint32_t omega, uint32_t round, uint32_t* part, uint32_t* psizes, uint32_t* applied);
/* LP refinement: up to `rounds` Jacobi rounds, ending at a round that applies nothing or
 * raises lambda-1 (undone). */
int orc_lp_refine(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint64_t* toff,
                  const uint32_t* tch, const uint32_t* inb, const float* w, const uint32_t* sizes, uint32_t k,
                  uint32_t omega, uint32_t rounds, uint32_t* part, uint32_t* psizes);
/* Best-fit-decreasing start of the device initial partitioner (seeded shuffle, sizes desc). */
int orc_bfd_start(uint32_t n, const uint32_t* sizes, uint32_t k, uint32_t omega, uint64_t seed, uint32_t* part,
                  uint32_t* loads);
/* The device initial partitioner restated (merged level, further coarsening, multistart
 * BFD + Jacobi, projection with k-way refinement). */
int orc_device_initial_partition(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const float* w,
                                 const uint32_t* sizes, uint32_t k, uint32_t omega, int prune, uint32_t starts,
                                 uint32_t rounds, uint32_t per_part, uint32_t repeats, uint32_t* part,
                                 uint32_t* psizes);

/* getPartitionsHypergraph(part, 2, true) (includes/hgraph.hpp:286-376) with exact set keys,
 * merged hedges in order of first appearance; two-phase (null arrays: sizes only); 1 when
 * the partition ids are not 0..k-1, all used. */
int orc_quotient(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                 const float* w, const uint32_t* part, uint32_t* k_out, uint32_t* e_out, uint64_t* p_out,
                 uint64_t* qoff, uint32_t* qpins, uint32_t* qsrcc, float* qw);

/* Whole pipeline (sources/main.cu:38-940).  ip_device: 0 the bisection, 1 the device initial
 * partitioner's restatement, 2 the device one on levels of at least 2^21 pins; ip_options: its (starts, rounds, nodes per part,
 * refinement rounds), NULL or 0 entries for the defaults (2, 8, 16, 4); lp_rounds: Jacobi
 * rounds before each level's FM rounds (k-way). on a host CSR; part[n] out.  Returns 0 on success. */
int orc_partition(int mode, uint32_t omega, uint32_t delta, uint32_t max_parts, uint32_t kway, float eps,
                  uint32_t cnc, uint32_t refine_repeats, int ipm, int prune, int ip_device,
                  const uint32_t* ip_options, uint32_t lp_rounds, uint32_t n, uint32_t e,
                  const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc, const float* w,
                  uint32_t* part, uint32_t* num_levels);

/* Validity (sources/constr.cpp:90-147): 1 valid, 0 invalid, -1 ids not sequential. */
int orc_check_validity(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins,
                       const uint32_t* srcc, const uint32_t* part, uint32_t omega, uint32_t delta,
                       uint32_t max_parts);

/* lambda-1 / cut-net / SOED over the rule-2 quotient hypergraph (includes/hgraph.hpp:214-376). */
int orc_metrics(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                const float* w, const uint32_t* part, double* km1, double* cutnet, double* soed);

#ifdef __cplusplus
}
#endif
#endif
