/*
 * hgp.h - C ABI of the MI355X-native multilevel hypergraph partitioner.
 *
 * Every entry point replaces one host step of the reference `hgraph_gpu`
 * pipeline (EMJzero/AxonCUDA).  The reference steps are C++ functions that take
 * raw device pointers and return callee-allocated device buffers that the caller
 * frees; the entry points below keep that contract with plain C types:
 *   - `d_` pointers are device (HBM) pointers, `h_` pointers are host pointers;
 *   - buffers returned through `T** d_out` are allocated by the library and are
 *     released with hgp_free();
 *   - buffers passed in as `T* d_out` are caller-allocated and sized as stated;
 *   - every call returns an hgp_status; no call exits or aborts the process.
 *     hgp_ctx_last_error() gives the message of the last failure.
 * Constraints (node cap, inbound cap) are per-call arguments: there is no
 * process-global device state, so concurrent jobs with different constraints
 * can share one context.  All calls on one context are serialised on the
 * context's stream (hgp_ctx_set_stream); a context is not thread-safe.
 *
 * Canonical orders (what the parity tests compare bit-exactly):
 *   touching   : per node, inbound hedges ascending, then outbound ascending.
 *   neighbours : per node, ascending node ids, the node itself excluded.
 *   coarse hedge: sources ascending, then destinations descending.
 */
#ifndef HGP_H
#define HGP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t hgp_status;
#define HGP_OK 0
#define HGP_E_INVALID 1    /* bad argument */
#define HGP_E_HIP 2        /* HIP runtime error */
#define HGP_E_NOMEM 3      /* device allocation failed / over the HBM budget */
#define HGP_E_CAPACITY 4   /* an internal table overflowed (retry with larger -om) */
#define HGP_E_IO 5         /* file could not be read or written / parse error */
#define HGP_E_INFEASIBLE 6 /* the constraints cannot be met */
#define HGP_E_NODEVICE 7   /* no usable gfx950 device */
#define HGP_E_INTERNAL 8

#define HGP_MODE_INCC 0 /* inbound-constrained minimal partitioning (-c / -m) */
#define HGP_MODE_KWAY 1 /* k-way epsilon-balanced partitioning (-k) */

typedef struct hgp_ctx hgp_ctx;
typedef struct hgp_hgraph hgp_hgraph;

/* Run configuration: the reference `runconfig` (includes/runconfig.hpp:32-51)
 * plus the resolved constraint triple (sources/runconfig.cpp:186-217). */
typedef struct hgp_params {
  int32_t mode;                     /* HGP_MODE_* */
  uint32_t max_nodes_per_part;      /* Omega */
  uint32_t max_inbound_per_part;    /* Delta */
  uint32_t max_parts;               /* in KWAY mode it must equal kway (HGP_E_INVALID otherwise) */
  uint32_t kway;                    /* -k k   (KWAY only) */
  float epsilon;                    /* -k eps (KWAY only) */
  uint32_t candidates_count;        /* -cnc, 1..16 (default 16) */
  uint32_t refine_repeats;          /* -rfr (default 16) */
  uint32_t save_memory_up_to_level; /* -smh (default 2) */
  float oversized_multiplier;       /* -om (default 1.5) */
  int32_t device_touching_construction; /* -dtc */
  int32_t initial_partitions_merge;     /* -ipm */
  int32_t verbose;                  /* -v 0..3 */
  /* k-way initial partitioner when Mt-KaHyPar is not installed: HGP_IP_HOST (recursive
   * bisection on the host), HGP_IP_DEVICE (the level coarsened further on the device,
   * best-fit-decreasing starts improved by label propagation, projected back with k-way
   * refinement) or HGP_IP_AUTO (the default: the device partitioner when the level holds at
   * least HGP_IP_AUTO_DEVICE_PINS pins, where the bisection's host time grows, the bisection
   * below).  No counterpart in the reference. */
  int32_t initial_partitioner;
  /* f > 0: before the k-way initial partition, drop hedges by the reference's degree-score
   * pruning and rescale the kept weights (sources/init_part.cu:749-817), aiming at f times
   * the reference's target of N * floor(P / E) kept hedges (f = 1 is the reference's rule);
   * 0: hand every hedge to the initial partitioner; -1 (default): 1 when Mt-KaHyPar runs
   * (the reference's behaviour), 0 for the built-in partitioners. */
  int32_t initial_prune;
  /* k-way only: up to this many Jacobi label-propagation rounds before each level's FM
   * rounds (0, the default: none).  A round that applies nothing ends them; a round that
   * raises lambda-1 is undone and ends them.  The reference's only label propagation is its
   * legacy initial partitioner (kernels/init_part.cu:89-151). */
  uint32_t lp_rounds;
  /* The device initial partitioner (HGP_IP_DEVICE): BFD starts, Jacobi rounds per start,
   * coarsening target in nodes per part, and k-way refinement rounds per projected level
   * (0 in any of them: the built-in default). */
  uint32_t ip_starts;
  uint32_t ip_rounds;
  uint32_t ip_nodes_per_part;
  uint32_t ip_refine_repeats;
  /* INCC refinement: device bytes its inbound-set event stage may take at once; the events of
   * a round that need more are processed in ranges of partitions, one after the other, with
   * the same result (0, the default: half of the device memory free at that moment).  The
   * reference sizes this stage without a bound (sources/refinement.cu:307-491). */
  uint64_t refine_event_budget;
} hgp_params;

enum { HGP_IP_HOST = 0, HGP_IP_DEVICE = 1, HGP_IP_AUTO = 2 };
#define HGP_IP_AUTO_DEVICE_PINS 2097152u

/* Launch parameters chosen by the autotuner.  No counterpart in the reference, whose block
 * sizes, hash-set sizes and LDS budgets are compile-time constants (headers/construction.cuh:
 * 22-29, headers/coarsening.cuh:21-31).  They change speed only: every entry point returns
 * the same result under any valid setting. */
typedef struct hgp_tuning {
  /* radix sorts of 2^c .. 2^(c+1) - 1 keys: keys per thread in a workgroup's tile (the tile is
   * 256 x this many keys), 1..16; 0 = the built-in rule (about 512 tiles per sort) */
  uint32_t sort_rounds[40];
  /* candidate scoring of heavy nodes: slots of the first-attempt HBM hash table (a power of
   * two, 2^13..2^24) and its fill limit in sixteenths (4..14), i.e. the table occupancy that
   * the reference's -om multiplier sets for its spill hash sets */
  uint32_t heavy_table_slots;
  uint32_t heavy_table_fill16;
  /* threads per workgroup of the thread-per-hedge cascade: 64, 128, 256, 512 or 1024 */
  uint32_t cascade_block;
  /* scans of more than this many 4096-element tiles take three passes instead of the
   * single-pass look-back */
  uint32_t scan_lookback_tiles;
  /* k-way gains from partition masks (k <= 64): the mean touching hedges per node of a level
   * from which each node takes 4 waves, and from which it takes 16, instead of one */
  uint32_t mask_wide4_degree;
  uint32_t mask_wide16_degree;
  /* independent kernels of one stage on side streams (default priority) beside the context's
   * stream: bit 0 the candidate tiers, bit 1 the cascade's sorted hedge classes beside the
   * size and inbound events, bit 2 (with bit 1) beside the small-hedge cascade too */
  uint32_t side_streams;
  /* radix sorts of more than 16,384 and at most coop_max_n keys run all their passes in one
   * launch of at most 64 workgroups that meet at grid barriers, each taking at least
   * coop_tile keys (0 disables them; serving workers never use them) */
  uint32_t coop_max_n;
  uint32_t coop_tile;
  /* 1 when set by hgp_autotune */
  uint32_t tuned;
} hgp_tuning;

/* Per-run statistics filled by hgp_partition. */
typedef struct hgp_run_stats {
  double core_ms;       /* device-resident input -> partition on host */
  double device_ms;     /* upload included */
  double ip_ms;         /* initial partitioning (KWAY) */
  uint32_t levels;      /* coarsening levels executed */
  uint32_t num_partitions;
  uint64_t total_pins;
  uint64_t bytes_algorithmic; /* compulsory traffic model, summed over levels */
  uint32_t level_nodes[64];
  uint64_t level_pins[64];
  uint64_t peak_device_bytes; /* most device memory the run held at once (size classes) */
} hgp_run_stats;

/* ---- context and device memory ------------------------------------------ */
void hgp_default_params(hgp_params* p);
hgp_status hgp_ctx_create(int device, uint64_t hbm_budget, hgp_ctx** out);
void hgp_ctx_destroy(hgp_ctx* ctx);
const char* hgp_ctx_last_error(const hgp_ctx* ctx);
hgp_status hgp_ctx_set_stream(hgp_ctx* ctx, void* hip_stream);
hgp_status hgp_synchronize(hgp_ctx* ctx);
hgp_status hgp_malloc(hgp_ctx* ctx, uint64_t bytes, void** d_ptr);
hgp_status hgp_free(hgp_ctx* ctx, void* d_ptr);
hgp_status hgp_memcpy_htod(hgp_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes);
hgp_status hgp_memcpy_dtoh(hgp_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes);
/* Device-to-device copy on the context stream (synchronous); no counterpart in the reference. */
hgp_status hgp_memcpy_dtod(hgp_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes);
/* Device memory held by the context: bytes handed out, bytes of freed blocks kept for reuse,
 * and the most handed out at once.  hgp_ctx_trim returns the kept blocks to HIP.  No
 * counterpart in the reference (it frees every buffer with cudaFree). */
hgp_status hgp_ctx_memory(const hgp_ctx* ctx, uint64_t* in_use, uint64_t* cached, uint64_t* peak);
hgp_status hgp_ctx_trim(hgp_ctx* ctx);

/* ---- host model: loaders and savers (includes/hgraph.hpp) ---------------- */
/* Loads .hgr (hMETIS, loadhMETIS :936-1013), .snn (loadSNN :707-743) or .axh
 * (loadAXH :805-854) by extension, then removes duplicate pins per hedge keeping
 * sources (deduplicateHyperedges(2,false) :518-581, called at sources/main.cu:125). */
hgp_status hgp_hgraph_load(const char* path, int dedup, hgp_hgraph** out);
hgp_status hgp_hgraph_from_csr(uint32_t num_nodes, uint32_t num_hedges, const uint64_t* h_offsets,
                               const uint32_t* h_pins, const uint32_t* h_srcs_count,
                               const float* h_weights, int dedup, hgp_hgraph** out);
void hgp_hgraph_free(hgp_hgraph* hg);
void hgp_hgraph_sizes(const hgp_hgraph* hg, uint32_t* num_nodes, uint32_t* num_hedges, uint64_t* num_pins);
hgp_status hgp_hgraph_copy_out(const hgp_hgraph* hg, uint64_t* h_offsets, uint32_t* h_pins,
                               uint32_t* h_srcs_count, float* h_weights);
/* Partition file: one id per line (savePartitioning :1054-1059). */
hgp_status hgp_save_partitioning(const char* path, const uint32_t* h_part, uint32_t num_nodes);
/* Quotient hypergraph writer for -s (sources/runconfig.cpp:219-247):
 * getPartitionsHypergraph(part,2,true) (:286-376) saved as .hgr/.snn/.axh. */
hgp_status hgp_save_partitioned(const hgp_hgraph* hg, const uint32_t* h_part, const char* path);
/* Metrics on the host model (lambda-1 :214-222 over the quotient hypergraph, cut-net, SOED). */
hgp_status hgp_hgraph_metrics(const hgp_hgraph* hg, const uint32_t* h_part, double* connectivity,
                              double* cutnet, double* soed);
/* Validity: count, node cap, distinct inbound per partition (sources/constr.cpp:90-147).
 * *valid = 1 when valid. */
hgp_status hgp_check_validity(const hgp_hgraph* hg, const uint32_t* h_part, uint32_t max_nodes_per_part,
                              uint32_t max_inbound_per_part, uint32_t max_parts, int32_t* valid);

/* ---- synthetic workloads (seeded, deterministic) ------------------------- */
/* Random hypergraph: |e| = 2 + Geometric(1/2) capped at max_size, uniform distinct pins. */
hgp_status hgp_gen_random(uint32_t num_nodes, uint32_t num_hedges, uint32_t max_size, uint64_t seed,
                          hgp_hgraph** out);
/* Netlist-shaped: power-law sizes (alpha, [2,max_size]), 2D-grid locality (frac_local of the
 * pins within `radius` of the first pin), the rest uniform. */
hgp_status hgp_gen_netlist(uint32_t num_nodes, uint32_t num_hedges, double alpha, uint32_t max_size,
                           uint32_t radius, double frac_local, uint64_t seed, hgp_hgraph** out);
/* Layered feed-forward SNN: `layers` x `per_layer` neurons on a grid of side `grid`; every
 * neuron of layers 0..L-2 is one hedge (src = itself) with `fanout` distinct destinations in
 * the next layer drawn from a 2D Gaussian of std `sigma`. */
hgp_status hgp_gen_snn(uint32_t layers, uint32_t per_layer, uint32_t grid, uint32_t fanout, double sigma,
                       uint64_t seed, hgp_hgraph** out);
hgp_status hgp_hgraph_save(const hgp_hgraph* hg, const char* path);

/* ---- stages (device pointers) -------------------------------------------- */

/* buildTouching (sources/construction.cu:64-179; kernels/construction.cu:9-84).
 * Incidence CSR: touching_size == number of pins. */
hgp_status hgp_build_touching(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                              const uint32_t* d_srcs_count, uint32_t num_nodes, uint32_t num_hedges,
                              uint64_t* touching_size, uint32_t** d_touching,
                              uint64_t** d_touching_offsets, uint32_t** d_inbound_count);

/* buildNeighbors (sources/construction.cu:242-340; kernels/construction.cu:157-297).
 * N(v) = union of the pins of v's touching hedges, minus v. */
hgp_status hgp_build_neighbors(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                               const uint32_t* d_touching, const uint64_t* d_touching_offsets,
                               uint32_t num_nodes, uint64_t* neighbors_size, uint32_t** d_neighbors,
                               uint64_t** d_neighbors_offsets);

/* candidatesProposal (sources/coarsening.cu:17-68; kernels/coarsening.cu:10-213).
 * Writes d_pairs/d_scores [num_nodes * candidates_count]: the top candidates by
 * (score desc, id desc); empty slots are UINT32_MAX / 0.  Scores are computed on the
 * fly from the hypergraph (no neighbour sets needed), so the reference's
 * d_neighbors arguments have no counterpart here. */
hgp_status hgp_candidates(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                          const uint32_t* d_srcs_count, const uint32_t* d_touching,
                          const uint64_t* d_touching_offsets, const uint32_t* d_inbound_count,
                          const float* d_hedge_weights, const uint32_t* d_nodes_sizes, uint32_t num_nodes,
                          uint32_t num_hedges, uint32_t candidates_count, uint32_t max_nodes_per_part,
                          uint32_t max_inbound_per_part, uint32_t* d_pairs, uint32_t* d_scores);

/* groupNodes (sources/coarsening.cu:70-210; kernels/coarsening.cu:241-505;
 * build_orphan_pairs sources/chaining.cu:260-329).  Pairs nodes along their
 * candidates (groups of at most two), pairs leftover candidate-less nodes, and
 * relabels groups zero-based in ascending raw-id order. */
hgp_status hgp_group_nodes(hgp_ctx* ctx, const uint32_t* d_inbound_count, const uint32_t* d_pairs,
                           const uint32_t* d_scores, const uint32_t* d_nodes_sizes, uint32_t num_nodes,
                           uint32_t candidates_count, uint32_t max_nodes_per_part,
                           uint32_t max_inbound_per_part, uint32_t* new_num_nodes, uint32_t** d_groups,
                           uint32_t** d_groups_sizes, uint32_t** d_ungroups, uint64_t** d_ungroups_offsets);

/* coarsenNeighbors (sources/construction.cu:342-561):
 * N'(g) = { groups[x] : x in N(m), m in g } minus g. */
hgp_status hgp_coarsen_neighbors(hgp_ctx* ctx, const uint32_t* d_neighbors,
                                 const uint64_t* d_neighbors_offsets, const uint32_t* d_groups,
                                 const uint32_t* d_ungroups, const uint64_t* d_ungroups_offsets,
                                 uint32_t num_nodes, uint32_t new_num_nodes, uint64_t* new_neighbors_size,
                                 uint32_t** d_coarse_neighbors, uint64_t** d_coarse_neighbors_offsets);

/* coarsenHedges (sources/construction.cu:563-676; kernels/construction.cu:304-489).
 * Coarse hedge = distinct groups of its pins; a source group that is also a
 * destination group is dropped from the sources; the hedge count never shrinks. */
hgp_status hgp_coarsen_hedges(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                              const uint32_t* d_srcs_count, const uint32_t* d_groups, uint32_t num_hedges,
                              uint64_t hedges_size, uint64_t* new_hedges_size, uint32_t** d_coarse_hedges,
                              uint64_t** d_coarse_hedges_offsets, uint32_t** d_coarse_srcs_count);

/* coarsenTouching (sources/construction.cu:678-789).  Built here as the incidence
 * of the coarse hedges, which equals the reference's union of the members'
 * inbound sets followed by the members' outbound sets minus the inbound ones. */
hgp_status hgp_coarsen_touching(hgp_ctx* ctx, const uint32_t* d_coarse_hedges,
                                const uint64_t* d_coarse_hedges_offsets, const uint32_t* d_coarse_srcs_count,
                                uint32_t new_num_nodes, uint32_t num_hedges, uint64_t* new_touching_size,
                                uint32_t** d_coarse_touching, uint64_t** d_coarse_touching_offsets,
                                uint32_t** d_coarse_inbound_count);

/* apply_uncoarsening_partitions (kernels/construction.cu:953-968). */
hgp_status hgp_uncoarsen(hgp_ctx* ctx, const uint32_t* d_groups, const uint32_t* d_coarse_partitions,
                         uint32_t num_nodes, uint32_t* d_partitions);

/* refinementRepeats (sources/refinement.cu:18-582): `refine_repeats` rounds of
 * parallel FM with size and inbound-set event checks.  Updates d_partitions and
 * d_partitions_sizes in place; when update_final_inbound_counts is set the exact
 * inbound-set size of every partition is written to d_partitions_inbound_sizes. */
hgp_status hgp_refine(hgp_ctx* ctx, const hgp_params* params, const uint32_t* d_hedges,
                      const uint64_t* d_hedges_offsets, const uint32_t* d_srcs_count,
                      const uint32_t* d_touching, const uint64_t* d_touching_offsets,
                      const uint32_t* d_inbound_count, const float* d_hedge_weights,
                      const uint32_t* d_nodes_sizes, uint32_t level_idx, uint32_t num_nodes,
                      uint32_t num_hedges, uint32_t num_partitions, int32_t update_final_inbound_counts,
                      uint32_t* d_partitions, uint32_t* d_partitions_sizes,
                      uint32_t* d_partitions_inbound_sizes);

/* Label-propagation refinement, k-way (modelled on the reference's legacy Jacobi rounds,
 * kernels/init_part.cu:89-151): up to params->lp_rounds rounds in which every node proposes
 * its best strictly-fitting move (tie hash mix32(q + round)) and the positive proposals are
 * applied by (target, gain desc, node) while each target's size before the round plus what
 * entered it fits max_nodes_per_part; a round that applies nothing ends them, and a round
 * that raises lambda-1 is undone and ends them.  Updates d_partitions and
 * d_partitions_sizes in place. */
hgp_status hgp_refine_lp(hgp_ctx* ctx, const hgp_params* params, const uint32_t* d_hedges,
                         const uint64_t* d_hedges_offsets, const uint32_t* d_srcs_count,
                         const uint32_t* d_touching, const uint64_t* d_touching_offsets,
                         const uint32_t* d_inbound_count, const float* d_hedge_weights,
                         const uint32_t* d_nodes_sizes, uint32_t num_nodes, uint32_t num_hedges,
                         uint32_t num_partitions, uint32_t* d_partitions, uint32_t* d_partitions_sizes);

/* One analysis round of refinementRepeats without applying it, exposing every
 * intermediate (host outputs, each num_nodes long):
 *   moves[v], gains[v]   : fm_refinement_gains_kernel (kernels/refinement.cu:82-209),
 *                          gains in 2^18 fixed point;
 *   ranks[v]             : chaining (sources/chaining.cu:19-254) or gain sort (:172-189);
 *   seq_gains[r]         : fm_refinement_cascade_kernel (:216-322) + inclusive scan;
 *   size_viol[r], inbound_viol[r]: scanned event flags (:239-305, :307-491);
 *   *best_rank           : best_move_functor max_element (:493-500). */
hgp_status hgp_refine_analyze(hgp_ctx* ctx, const hgp_params* params, const uint32_t* d_hedges,
                              const uint64_t* d_hedges_offsets, const uint32_t* d_srcs_count,
                              const uint32_t* d_touching, const uint64_t* d_touching_offsets,
                              const uint32_t* d_inbound_count, const float* d_hedge_weights,
                              const uint32_t* d_nodes_sizes, uint32_t num_nodes, uint32_t num_hedges,
                              uint32_t num_partitions, uint32_t fm_repeat, int32_t chainup,
                              const uint32_t* d_partitions, const uint32_t* d_partitions_sizes,
                              uint32_t* h_moves, int64_t* h_gains, uint32_t* h_ranks, int64_t* h_seq_gains,
                              int32_t* h_size_viol, int32_t* h_inbound_viol, uint32_t* h_best_rank);

/* greedyMergeGroups (-ipm, sources/postprocess.cu:16-245). */
hgp_status hgp_greedy_merge_groups(hgp_ctx* ctx, const uint32_t* d_nodes_sizes,
                                   const uint32_t* d_inbound_count, uint32_t num_nodes, uint32_t num_groups,
                                   uint32_t max_nodes_per_part, uint32_t max_inbound_per_part,
                                   uint32_t* d_groups, uint32_t* d_groups_sizes, uint32_t* new_num_groups);

/* mergeSmallPartitions (INCC, sources/postprocess.cu:247-420). */
hgp_status hgp_merge_small_partitions(hgp_ctx* ctx, const uint32_t* d_partitions_sizes,
                                      const uint32_t* d_partitions_inbound_sizes, uint32_t num_nodes,
                                      uint32_t num_partitions, uint32_t max_nodes_per_part,
                                      uint32_t max_inbound_per_part, uint32_t* d_partitions);

/* Final zero-based relabel in ascending old-id order (sources/main.cu:817-830). */
hgp_status hgp_relabel_partitions(hgp_ctx* ctx, uint32_t num_nodes, uint32_t* d_partitions,
                                  uint32_t* new_num_partitions);

/* Exact inbound-set size per partition (inbound_sets_size_kernel,
 * kernels/refinement.cu:1176-1219): d_out[num_partitions]. */
hgp_status hgp_inbound_sizes(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                             const uint32_t* d_srcs_count, const uint32_t* d_partitions, uint32_t num_hedges,
                             uint32_t num_partitions, uint32_t* d_out);

/* getPartitionsHypergraph(part, 2, true) (includes/hgraph.hpp:286-376) on the device: the
 * partitioned hypergraph, one node per partition.  Every hedge becomes (its distinct source
 * partitions, its distinct destination partitions that are not source ones); hedges with at
 * most one source partition and no destination partition are dropped; identical hedges are
 * merged with their weights summed.  Unlike the reference (a 64-bit hash of the sets as a map
 * key, unordered output), equal sets are compared exactly and the merged hedges come in order
 * of first appearance, sources then destinations, each ascending, each weight summed in float
 * in hedge order: the same hypergraph hgp_save_partitioned writes.  Partition ids must be
 * 0..k-1 with every id used (HGP_E_INVALID otherwise, as the reference throws).  Allocates
 * *d_pins[*num_pins], *d_offsets[*num_merged + 1] (u64), *d_srcs_count[*num_merged] and
 * *d_weights[*num_merged]; *num_partitions = k. */
hgp_status hgp_quotient(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                        const uint32_t* d_srcs_count, const float* d_hedge_weights, const uint32_t* d_partitions,
                        uint32_t num_nodes, uint32_t num_hedges, uint32_t* num_partitions, uint32_t* num_merged,
                        uint64_t* num_pins, uint32_t** d_pins, uint64_t** d_offsets, uint32_t** d_srcs_count_out,
                        float** d_weights);

/* lambda-1 on the device: sum_e w_e * (lambda(e) - 1), in 2^18 fixed point
 * (compute_connectivity_kernel, kernels/init_part.cu:287-314). */
hgp_status hgp_connectivity(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                            const float* d_hedge_weights, const uint32_t* d_partitions, uint32_t num_hedges,
                            int64_t* km1_fixed);

/* The autotuner (no counterpart in the reference): times every hgp_tuning knob's candidates
 * on synthetic inputs built on this context's device (sorts of 2^15 .. 2^max_log2 keys,
 * max_log2 = 0 for 26; about a few seconds) and keeps the fastest in the context; *out (may
 * be NULL) receives the result.  Results of every entry point are unchanged by it. */
hgp_status hgp_autotune(hgp_ctx* ctx, uint32_t max_log2, hgp_tuning* out);
/* The context's current launch parameters, and replacing them (HGP_E_INVALID when a value is
 * out of its range; nothing is changed then). */
hgp_status hgp_tuning_get(const hgp_ctx* ctx, hgp_tuning* out);
hgp_status hgp_tuning_set(hgp_ctx* ctx, const hgp_tuning* in);
/* The built-in defaults, and the launch parameters that contexts created from now on start
 * with (serving workers included); NULL restores the built-in defaults. */
void hgp_tuning_defaults(hgp_tuning* out);
hgp_status hgp_tuning_set_process_default(const hgp_tuning* in);
/* A tuning as a text file ("key value" lines), and back (HGP_E_IO / HGP_E_INVALID on a
 * missing or malformed file). */
hgp_status hgp_tuning_save(const hgp_tuning* t, const char* path);
hgp_status hgp_tuning_load(const char* path, hgp_tuning* out);

/* hgp_connectivity without any device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The result goes to device memory:
 * *d_km1_fixed = the same 2^18 fixed-point lambda-1, or -1 when a pin's partition id is
 * >= num_partitions.  Launches two kernels on the context stream (the call returns before
 * they finish) and calls no hipMalloc, hipFree or memset.  Its partial sums, ticket and
 * error flag live in a fixed library array of 16.5 KiB per device (2048 workgroup
 * partials); num_partitions above 65536 is rejected with HGP_E_CAPACITY.  Calls that
 * overlap on one device share that array and race: serialise them (one stream, or events). */
hgp_status hgp_connectivity_noalloc(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                                    const float* d_hedge_weights, const uint32_t* d_partitions, uint32_t num_hedges,
                                    uint32_t num_partitions, int64_t* d_km1_fixed);

/* initial_partitioning_kahypar (sources/init_part.cu:733-953): k-way initial
 * partition of a coarse level.  With `prune` = f > 0, hedges are first pruned as the
 * reference prunes them (armonic_degree_score_kernel + prune_hedges_kernel,
 * kernels/init_part.cu:221-281) with a splitmix64 draw per hedge in place of cuRAND Philox,
 * aiming at f * N * floor(P / E) kept hedges (f = 1: the reference's target).  Runs `mtkahypar` when it is
 * on PATH (same command line as the reference); otherwise the built-in recursive-bisection
 * partitioner.  Allocates d_partitions[num_nodes] and d_partitions_sizes[k]. */
hgp_status hgp_initial_partition(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                                 const float* d_hedge_weights, uint32_t num_nodes, uint32_t num_hedges,
                                 uint64_t hedges_size, const uint32_t* d_nodes_sizes, uint32_t k, float epsilon,
                                 uint32_t max_nodes_per_part, int32_t prune, uint32_t** d_partitions,
                                 uint32_t** d_partitions_sizes);

/* The initial partitioner's input, built on the device exactly as hgp_initial_partition
 * builds it (sources/init_part.cu:749-888): optional pruning, then hedges of >= 2 distinct
 * pins with sorted pins, identical pin sets merged into their first appearance with their
 * 2^18 fixed-point weights summed.  Two-phase: with NULL arrays only *num_merged and
 * *merged_pins are set; then h_offsets[num_merged + 1], h_pins[merged_pins] and
 * h_weights[num_merged] receive the result.  For parity tests of that step. */
hgp_status hgp_ip_input(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                        const float* d_hedge_weights, const uint32_t* d_nodes_sizes, uint32_t num_nodes,
                        uint32_t num_hedges, int32_t prune, uint32_t* num_merged, uint64_t* merged_pins,
                        uint64_t* h_offsets, uint32_t* h_pins, uint64_t* h_weights);

/* ---- whole pipeline (sources/main.cu:38-940) ------------------------------ */
/* Partitions the host model; writes h_partitions_out[num_nodes] (zero-based ids). */
hgp_status hgp_partition(hgp_ctx* ctx, const hgp_params* params, const hgp_hgraph* hg,
                         uint32_t* h_partitions_out, hgp_run_stats* stats);

/* ---- kernel profiler (no counterpart in the reference) --------------------
 * Times every launch of one kernel (by name, e.g. "k_cand_light"; NULL or "" stops) with
 * HIP events on the context stream; hgp_profile_read synchronizes, returns the summed
 * device time, the launch count and the launches' algorithmic bytes (0 for kernels without
 * a byte model), and clears the record. */
hgp_status hgp_profile_kernel(hgp_ctx* ctx, const char* kernel_name);
/* on != 0: the profiled kernel's launches also count their algorithmic bytes on the device,
 * by a model that adds device work (for a separate, untimed run); 0 (default): timing only. */
hgp_status hgp_profile_count(hgp_ctx* ctx, int32_t on);
hgp_status hgp_profile_read(hgp_ctx* ctx, double* total_ms, uint64_t* launches, uint64_t* algorithmic_bytes);

/* Same pipeline on a hypergraph already resident in HBM (sources first, rule-2
 * deduplicated as hgp_hgraph_load leaves it).  The inputs are only read.  Contract:
 * d_hedges_offsets[0] == 0, offsets non-decreasing, every pin < num_nodes; the first and
 * the last are checked (HGP_E_INVALID), the order of the offsets is the caller's to keep. */
hgp_status hgp_partition_device(hgp_ctx* ctx, const hgp_params* params, uint32_t num_nodes, uint32_t num_hedges,
                                const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                                const uint32_t* d_srcs_count, const float* d_hedge_weights,
                                uint32_t* h_partitions_out, hgp_run_stats* stats);

/* ---- continuous-batching service (no counterpart in the reference) ---------
 * Runs submitted jobs concurrently on one GPU: `workers` host threads, each with its own
 * context and HIP stream, take queued jobs whose estimated HBM footprint fits next to the
 * running ones (hbm_budget 0 = 90% of the free device memory).  The hypergraph passed to
 * submit must stay alive until hgp_server_wait returns for that job.  Results equal those
 * of hgp_partition on the same input.  latency_ms = submit -> done, queue_ms = submit ->
 * start. */
typedef struct hgp_server hgp_server;
/* Device bytes the service reserves for one job when admitting it: an upper bound of the
 * run's peak_device_bytes for the pipeline's per-level buffers. */
uint64_t hgp_job_footprint(const hgp_hgraph* hg, const hgp_params* params);
hgp_status hgp_server_create(int device, uint64_t hbm_budget, uint32_t workers, hgp_server** out);
hgp_status hgp_server_submit(hgp_server* server, const hgp_hgraph* hg, const hgp_params* params, uint64_t* job_id);
hgp_status hgp_server_wait(hgp_server* server, uint64_t job_id, uint32_t* h_partitions_out, hgp_run_stats* stats,
                           double* latency_ms, double* queue_ms);
void hgp_server_destroy(hgp_server* server);

#ifdef __cplusplus
}
#endif

#endif /* HGP_H */
