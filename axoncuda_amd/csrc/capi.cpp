// capi.cpp - extern "C" boundary of libhgp (declared in include/hgp.h).
//
// Each wrapper validates its arguments, runs the stage on the context stream and turns
// every hgp::Error (or std::exception) into a status code with the message kept in the
// context; nothing here exits or aborts the process.
#include <hip/hip_runtime.h>

#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <unordered_map>

#include "common.hpp"
#include "host/driver.hpp"
#include "host/ip_core.hpp"
#include "host/model.hpp"
#include "prim.hpp"
#include "stages.hpp"

struct hgp_ctx {
  hgp::Ctx c;
};

namespace hgp {

namespace {

uint64_t size_class(uint64_t bytes) {
  if (bytes <= 4096) return 4096;
  int b = 63 - __builtin_clzll(bytes - 1);  // 2^b < bytes <= 2^(b+1)
  const uint64_t step = 1ull << (b - 1);    // four classes per power of two
  return (bytes + step - 1) / step * step;
}

}  // namespace

void* dmalloc(Ctx& c, uint64_t bytes) {
  if (bytes == 0) return nullptr;
  const uint64_t cls = size_class(bytes);
  if (c.budget && c.in_use + cls > c.budget)
    throw Error(HGP_E_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes exceeds the HBM budget");
  void* p = nullptr;
  auto it = c.free_lists.find(cls);
  if (it != c.free_lists.end() && !it->second.empty()) {
    p = it->second.back();
    it->second.pop_back();
    c.cached -= cls;
  } else {
    if (c.budget && c.in_use + c.cached + cls > c.budget) release_cached(c);
    hipError_t e = hipMallocAsync(&p, cls, c.stream);
    if ((e != hipSuccess || !p) && c.cached) {  // out of memory: give the parked blocks back and retry
      (void)hipGetLastError();
      release_cached(c);
      e = hipMallocAsync(&p, cls, c.stream);
    }
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      throw Error(HGP_E_NOMEM, "hipMallocAsync(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
    }
  }
  c.in_use += cls;
  c.peak = std::max(c.peak, c.in_use);
  c.live[p] = cls;
  return p;
}

void dfree(Ctx& c, void* p) {
  if (!p) return;
  auto it = c.live.find(p);
  if (it == c.live.end()) return;  // not ours
  const uint64_t cls = it->second;
  c.live.erase(it);
  c.in_use -= cls;
  c.free_lists[cls].push_back(p);
  c.cached += cls;
}

void release_cached(Ctx& c) {
  for (auto& kv : c.free_lists)
    for (void* p : kv.second) (void)hipFreeAsync(p, c.stream);
  c.free_lists.clear();
  c.cached = 0;
  (void)hipStreamSynchronize(c.stream);
}

}  // namespace hgp

using hgp::Ctx;

namespace {

// launch parameters that new contexts start with (hgp_tuning_set_process_default)
std::mutex g_tune_mu;
hgp::Tuning g_default_tuning;

void to_public(const hgp::Tuning& t, hgp_tuning* o) {
  std::memset(o, 0, sizeof *o);
  for (int i = 0; i < 40; ++i) o->sort_rounds[i] = t.sort_rounds[i];
  o->heavy_table_slots = t.heavy_slots;
  o->heavy_table_fill16 = t.heavy_fill16;
  o->cascade_block = t.cascade_block;
  o->scan_lookback_tiles = t.scan_lookback_tiles;
  o->mask_wide4_degree = t.mask_wide4_degree;
  o->mask_wide16_degree = t.mask_wide16_degree;
  o->side_streams = t.side_streams;
  o->coop_max_n = t.coop_max_n;
  o->coop_tile = t.coop_tile;
  o->tuned = t.tuned ? 1u : 0u;
}

bool from_public(const hgp_tuning* o, hgp::Tuning& t) {
  for (int i = 0; i < 40; ++i)
    if (o->sort_rounds[i] > 16) return false;
  const uint32_t s = o->heavy_table_slots, b = o->cascade_block;
  if (s < (1u << 13) || s > (1u << 24) || (s & (s - 1))) return false;
  if (o->heavy_table_fill16 < 4 || o->heavy_table_fill16 > 14) return false;
  if (b < 64 || b > 1024 || (b & (b - 1))) return false;
  if (o->coop_tile < 256 || o->coop_tile > (1u << 20)) return false;
  for (int i = 0; i < 40; ++i) t.sort_rounds[i] = o->sort_rounds[i];
  t.heavy_slots = s;
  t.heavy_fill16 = o->heavy_table_fill16;
  t.cascade_block = b;
  t.scan_lookback_tiles = o->scan_lookback_tiles;
  t.mask_wide4_degree = o->mask_wide4_degree;
  t.mask_wide16_degree = o->mask_wide16_degree;
  t.side_streams = o->side_streams & 7u;
  t.coop_max_n = o->coop_max_n;
  t.coop_tile = o->coop_tile;
  t.tuned = o->tuned != 0;
  return true;
}

template <typename F>
hgp_status guard(hgp_ctx* ctx, F&& f) {
  if (!ctx) return HGP_E_INVALID;
  try {
    f(ctx->c);
    return HGP_OK;
  } catch (const hgp::Error& e) {
    ctx->c.last_error = e.what();
    if (ctx->c.verbose) fprintf(stderr, "hgp error: %s\n", e.what());
    return e.status;
  } catch (const std::bad_alloc&) {
    ctx->c.last_error = "host allocation failed";
    return HGP_E_NOMEM;
  } catch (const std::exception& e) {
    ctx->c.last_error = e.what();
    return HGP_E_INTERNAL;
  }
}

}  // namespace

extern "C" {

void hgp_default_params(hgp_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->mode = HGP_MODE_INCC;
  p->max_nodes_per_part = 1024;  // loihi64 (sources/constr.cpp:195-202)
  p->max_inbound_per_part = 4096;
  p->max_parts = 4096;
  p->kway = 0;
  p->epsilon = 0.0f;
  p->candidates_count = 16;          // MAX_CANDIDATES (headers/defines.cuh:49)
  p->refine_repeats = 16;            // REFINE_REPEATS (:54)
  p->save_memory_up_to_level = 2;    // SAVE_MEMORY_UP_TO_LEVEL (:31)
  p->oversized_multiplier = 1.5f;    // OVERSIZED_SIZE_MULTIPLIER (:26)
  p->device_touching_construction = 0;
  p->initial_partitions_merge = 0;
  p->verbose = 0;
  p->initial_partitioner = HGP_IP_AUTO;
  p->initial_prune = -1;  // the reference's pruning when Mt-KaHyPar runs, none otherwise
  p->lp_rounds = 0;
}

hgp_status hgp_ctx_create(int device, uint64_t hbm_budget, hgp_ctx** out) {
  if (!out) return HGP_E_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return HGP_E_NODEVICE;
  }
  if (device < 0 || device >= count) return HGP_E_NODEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return HGP_E_NODEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return HGP_E_NODEVICE;
  hgp_ctx* ctx = new hgp_ctx();
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    ctx->c.tune = g_default_tuning;
  }
  ctx->c.device = device;
  ctx->c.budget = hbm_budget;
  ctx->c.num_cus = (uint32_t)prop.multiProcessorCount;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return HGP_E_HIP;
  }
  ctx->c.own_stream = true;
  try {
    hgp::noalloc_setup(device);  // this device's no-allocation scratch (device is current)
  } catch (const hgp::Error& e) {
    (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
    return e.status;
  }
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
    uint64_t thr = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
  }
  *out = ctx;
  return HGP_OK;
}

void hgp_ctx_destroy(hgp_ctx* ctx) {
  if (!ctx) return;
  (void)hipStreamSynchronize(ctx->c.stream);
  for (auto& r : ctx->c.prof) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto& kv : ctx->c.live) (void)hipFreeAsync(kv.first, ctx->c.stream);
  hgp::release_cached(ctx->c);
  (void)hipStreamSynchronize(ctx->c.stream);
  hgp::destroy_sides(ctx->c);
  if (ctx->c.own_stream) (void)hipStreamDestroy(ctx->c.stream);
  if (ctx->c.pinned) (void)hipHostFree(ctx->c.pinned);
  delete ctx;
}

const char* hgp_ctx_last_error(const hgp_ctx* ctx) { return ctx ? ctx->c.last_error.c_str() : "null context"; }

hgp_status hgp_ctx_set_stream(hgp_ctx* ctx, void* stream) {
  return guard(ctx, [&](Ctx& c) {
    HGP_HIP(hipStreamSynchronize(c.stream));
    if (c.own_stream) HGP_HIP(hipStreamDestroy(c.stream));
    c.stream = (hipStream_t)stream;
    c.own_stream = false;
  });
}

hgp_status hgp_synchronize(hgp_ctx* ctx) { return guard(ctx, [&](Ctx& c) { hgp::sync(c); }); }

hgp_status hgp_malloc(hgp_ctx* ctx, uint64_t bytes, void** d_ptr) {
  if (!d_ptr) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    *d_ptr = hgp::dmalloc(c, bytes);
    hgp::sync(c);
  });
}

hgp_status hgp_free(hgp_ctx* ctx, void* d_ptr) {
  return guard(ctx, [&](Ctx& c) {
    hgp::dfree(c, d_ptr);
    hgp::sync(c);
  });
}

hgp_status hgp_memcpy_htod(hgp_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes) {
  return guard(ctx, [&](Ctx& c) {
    hgp::copy_h2d(c, d_dst, h_src, bytes);
    hgp::sync(c);
  });
}

hgp_status hgp_memcpy_dtod(hgp_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes) {
  return guard(ctx, [&](Ctx& c) {
    hgp::copy_d2d(c, d_dst, d_src, bytes);
    hgp::sync(c);
  });
}

hgp_status hgp_ctx_memory(const hgp_ctx* ctx, uint64_t* in_use, uint64_t* cached, uint64_t* peak) {
  if (!ctx) return HGP_E_INVALID;
  if (in_use) *in_use = ctx->c.in_use;
  if (cached) *cached = ctx->c.cached;
  if (peak) *peak = ctx->c.peak;
  return HGP_OK;
}

hgp_status hgp_ctx_trim(hgp_ctx* ctx) { return guard(ctx, [&](Ctx& c) { hgp::release_cached(c); }); }

hgp_status hgp_memcpy_dtoh(hgp_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes) {
  return guard(ctx, [&](Ctx& c) { hgp::copy_d2h(c, h_dst, d_src, bytes); });
}

// ---- host model ---------------------------------------------------------------------------
hgp_status hgp_hgraph_load(const char* path, int dedup, hgp_hgraph** out) {
  if (!path || !out) return HGP_E_INVALID;
  *out = nullptr;
  try {
    hgp_hgraph* hg = new hgp_hgraph();
    std::string err = hgp::load_hgraph(path, *hg);
    if (!err.empty()) {
      fprintf(stderr, "Error loading file: %s\n", err.c_str());
      delete hg;
      return HGP_E_IO;
    }
    if (dedup) hgp::dedup_rule2(*hg);
    *out = hg;
    return HGP_OK;
  } catch (...) {
    return HGP_E_NOMEM;
  }
}

hgp_status hgp_hgraph_from_csr(uint32_t num_nodes, uint32_t num_hedges, const uint64_t* h_offsets, const uint32_t* h_pins,
                               const uint32_t* h_srcs_count, const float* h_weights, int dedup, hgp_hgraph** out) {
  if (!out || !h_offsets || (num_hedges && (!h_pins && h_offsets[num_hedges] > 0))) return HGP_E_INVALID;
  *out = nullptr;
  try {
    hgp_hgraph* hg = new hgp_hgraph();
    hg->n = num_nodes;
    hg->off.assign(h_offsets, h_offsets + num_hedges + 1);
    const uint64_t p = h_offsets[num_hedges] - h_offsets[0];
    hg->pins.assign(h_pins + h_offsets[0], h_pins + h_offsets[0] + p);
    for (auto& o : hg->off) o -= h_offsets[0];
    if (h_srcs_count) hg->srcc.assign(h_srcs_count, h_srcs_count + num_hedges);
    else hg->srcc.assign(num_hedges, 0);
    if (h_weights) hg->w.assign(h_weights, h_weights + num_hedges);
    else hg->w.assign(num_hedges, 1.0f);
    for (uint32_t v : hg->pins)
      if (v >= num_nodes) {
        delete hg;
        return HGP_E_INVALID;
      }
    for (uint32_t i = 0; i < num_hedges; ++i)
      if (hg->off[i + 1] < hg->off[i] || hg->srcc[i] > hg->off[i + 1] - hg->off[i]) {
        delete hg;
        return HGP_E_INVALID;
      }
    if (dedup) hgp::dedup_rule2(*hg);
    *out = hg;
    return HGP_OK;
  } catch (...) {
    return HGP_E_NOMEM;
  }
}

void hgp_hgraph_free(hgp_hgraph* hg) { delete hg; }

void hgp_hgraph_sizes(const hgp_hgraph* hg, uint32_t* n, uint32_t* e, uint64_t* p) {
  if (!hg) return;
  if (n) *n = hg->n;
  if (e) *e = hg->e();
  if (p) *p = hg->pins.size();
}

hgp_status hgp_hgraph_copy_out(const hgp_hgraph* hg, uint64_t* off, uint32_t* pins, uint32_t* srcc, float* w) {
  if (!hg) return HGP_E_INVALID;
  if (off) std::memcpy(off, hg->off.data(), hg->off.size() * 8);
  if (pins) std::memcpy(pins, hg->pins.data(), hg->pins.size() * 4);
  if (srcc) std::memcpy(srcc, hg->srcc.data(), hg->srcc.size() * 4);
  if (w) std::memcpy(w, hg->w.data(), hg->w.size() * 4);
  return HGP_OK;
}

hgp_status hgp_save_partitioning(const char* path, const uint32_t* part, uint32_t n) {
  if (!path || (!part && n)) return HGP_E_INVALID;
  FILE* f = std::fopen(path, "w");
  if (!f) return HGP_E_IO;
  std::string buf;
  buf.resize((size_t)n * 11);  // at most 10 digits and a newline per node
  char* o = buf.data();
  for (uint32_t v = 0; v < n; ++v) {
    char tmp[10];
    uint32_t x = part[v];
    int l = 0;
    do {
      tmp[l++] = (char)('0' + x % 10);
      x /= 10;
    } while (x);
    while (l) *o++ = tmp[--l];
    *o++ = '\n';
  }
  std::fwrite(buf.data(), 1, (size_t)(o - buf.data()), f);
  std::fclose(f);
  return HGP_OK;
}

hgp_status hgp_save_partitioned(const hgp_hgraph* hg, const uint32_t* part, const char* path) {
  if (!hg || !part || !path) return HGP_E_INVALID;
  std::vector<uint32_t> p(part, part + hg->n);
  // the quotient needs partition ids 0..N-1 with none missing (includes/hgraph.hpp:292-295)
  uint32_t k = 0;
  for (uint32_t x : p) k = std::max(k, x + 1);
  std::vector<uint8_t> seen(k, 0);
  for (uint32_t x : p) seen[x] = 1;
  for (uint8_t s : seen)
    if (!s) return HGP_E_INVALID;
  hgp_hgraph q = hgp::partitions_hypergraph(*hg, p);
  // every quotient hedge needs a destination (the reference's HyperGraph constructor,
  // includes/hgraph.hpp:145, refuses the quotient otherwise)
  for (uint32_t i = 0; i < q.e(); ++i)
    if (q.off[i + 1] - q.off[i] == q.srcc[i]) return HGP_E_INVALID;
  return hgp::save_hgraph(q, path).empty() ? HGP_OK : HGP_E_IO;
}

hgp_status hgp_hgraph_save(const hgp_hgraph* hg, const char* path) {
  if (!hg || !path) return HGP_E_INVALID;
  return hgp::save_hgraph(*hg, path).empty() ? HGP_OK : HGP_E_IO;
}

hgp_status hgp_hgraph_metrics(const hgp_hgraph* hg, const uint32_t* part, double* km1, double* cutnet, double* soed) {
  if (!hg || !part) return HGP_E_INVALID;
  std::vector<uint32_t> p(part, part + hg->n);
  hgp::Metrics m = hgp::metrics(*hg, p);
  if (km1) *km1 = m.connectivity;
  if (cutnet) *cutnet = m.cutnet;
  if (soed) *soed = m.soed;
  return HGP_OK;
}

hgp_status hgp_check_validity(const hgp_hgraph* hg, const uint32_t* part, uint32_t omega, uint32_t delta,
                              uint32_t max_parts, int32_t* valid) {
  if (!hg || !part || !valid) return HGP_E_INVALID;
  std::vector<uint32_t> p(part, part + hg->n);
  *valid = hgp::validity(*hg, p, omega, delta, max_parts, nullptr);
  return HGP_OK;
}

hgp_status hgp_gen_random(uint32_t n, uint32_t e, uint32_t max_size, uint64_t seed, hgp_hgraph** out) {
  if (!out || n < 2 || max_size < 2) return HGP_E_INVALID;
  hgp_hgraph* hg = new hgp_hgraph();
  hgp::gen_random(n, e, max_size, seed, *hg);
  *out = hg;
  return HGP_OK;
}

hgp_status hgp_gen_netlist(uint32_t n, uint32_t e, double alpha, uint32_t max_size, uint32_t radius, double frac_local,
                           uint64_t seed, hgp_hgraph** out) {
  if (!out || n < 2 || max_size < 2) return HGP_E_INVALID;
  hgp_hgraph* hg = new hgp_hgraph();
  hgp::gen_netlist(n, e, alpha, max_size, radius, frac_local, seed, *hg);
  *out = hg;
  return HGP_OK;
}

hgp_status hgp_gen_snn(uint32_t layers, uint32_t per_layer, uint32_t grid, uint32_t fanout, double sigma, uint64_t seed,
                       hgp_hgraph** out) {
  if (!out || layers < 2 || per_layer == 0 || (uint64_t)grid * grid < per_layer || fanout >= per_layer) return HGP_E_INVALID;
  hgp_hgraph* hg = new hgp_hgraph();
  hgp::gen_snn(layers, per_layer, grid, fanout, sigma, seed, *hg);
  *out = hg;
  return HGP_OK;
}

// ---- stages -------------------------------------------------------------------------------
hgp_status hgp_build_touching(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                              const uint32_t* d_srcs_count, uint32_t num_nodes, uint32_t num_hedges,
                              uint64_t* touching_size, uint32_t** d_touching, uint64_t** d_touching_offsets,
                              uint32_t** d_inbound_count) {
  if (!touching_size || !d_touching || !d_touching_offsets || !d_inbound_count) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp::build_touching(c, d_hedges, d_hedges_offsets, d_srcs_count, num_nodes, num_hedges, *touching_size, *d_touching,
                        *d_touching_offsets, *d_inbound_count);
    hgp::sync(c);
  });
}

hgp_status hgp_build_neighbors(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                               const uint32_t* d_touching, const uint64_t* d_touching_offsets, uint32_t num_nodes,
                               uint64_t* neighbors_size, uint32_t** d_neighbors, uint64_t** d_neighbors_offsets) {
  if (!neighbors_size || !d_neighbors || !d_neighbors_offsets) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp::build_neighbors(c, d_hedges, d_hedges_offsets, d_touching, d_touching_offsets, num_nodes, *neighbors_size,
                         *d_neighbors, *d_neighbors_offsets);
    hgp::sync(c);
  });
}

hgp_status hgp_candidates(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                          const uint32_t* d_srcs_count, const uint32_t* d_touching, const uint64_t* d_touching_offsets,
                          const uint32_t* d_inbound_count, const float* d_hedge_weights, const uint32_t* d_nodes_sizes,
                          uint32_t num_nodes, uint32_t num_hedges, uint32_t candidates_count, uint32_t max_nodes_per_part,
                          uint32_t max_inbound_per_part, uint32_t* d_pairs, uint32_t* d_scores) {
  if (!d_pairs || !d_scores || candidates_count < 1 || candidates_count > 16) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp::candidates(c, d_hedges, d_hedges_offsets, d_srcs_count, d_touching, d_touching_offsets, d_inbound_count,
                    d_hedge_weights, d_nodes_sizes, num_nodes, num_hedges, candidates_count, max_nodes_per_part,
                    max_inbound_per_part, d_pairs, d_scores);
    hgp::sync(c);
  });
}

hgp_status hgp_group_nodes(hgp_ctx* ctx, const uint32_t* d_inbound_count, const uint32_t* d_pairs,
                           const uint32_t* d_scores, const uint32_t* d_nodes_sizes, uint32_t num_nodes,
                           uint32_t candidates_count, uint32_t max_nodes_per_part, uint32_t max_inbound_per_part,
                           uint32_t* new_num_nodes, uint32_t** d_groups, uint32_t** d_groups_sizes, uint32_t** d_ungroups,
                           uint64_t** d_ungroups_offsets) {
  if (!new_num_nodes || !d_groups || !d_groups_sizes || candidates_count < 1 || candidates_count > 16)
    return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp::group_nodes(c, d_inbound_count, d_pairs, d_scores, d_nodes_sizes, num_nodes, candidates_count,
                     max_nodes_per_part, max_inbound_per_part, *new_num_nodes, *d_groups, *d_groups_sizes, d_ungroups,
                     d_ungroups_offsets);
    hgp::sync(c);
  });
}

hgp_status hgp_coarsen_neighbors(hgp_ctx* ctx, const uint32_t* d_neighbors, const uint64_t* d_neighbors_offsets,
                                 const uint32_t* d_groups, const uint32_t* d_ungroups, const uint64_t* d_ungroups_offsets,
                                 uint32_t num_nodes, uint32_t new_num_nodes, uint64_t* new_neighbors_size,
                                 uint32_t** d_coarse_neighbors, uint64_t** d_coarse_neighbors_offsets) {
  (void)d_ungroups;
  (void)d_ungroups_offsets;
  if (!new_neighbors_size || !d_coarse_neighbors || !d_coarse_neighbors_offsets) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp::coarsen_neighbors(c, d_neighbors, d_neighbors_offsets, d_groups, num_nodes, new_num_nodes, *new_neighbors_size,
                           *d_coarse_neighbors, *d_coarse_neighbors_offsets);
    hgp::sync(c);
  });
}

hgp_status hgp_coarsen_hedges(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                              const uint32_t* d_srcs_count, const uint32_t* d_groups, uint32_t num_hedges,
                              uint64_t hedges_size, uint64_t* new_hedges_size, uint32_t** d_coarse_hedges,
                              uint64_t** d_coarse_hedges_offsets, uint32_t** d_coarse_srcs_count) {
  if (!new_hedges_size || !d_coarse_hedges || !d_coarse_hedges_offsets || !d_coarse_srcs_count) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    // new_n = 0: the coarse node count is derived on the device from the groups of the pins
    hgp::coarsen_hedges(c, d_hedges, d_hedges_offsets, d_srcs_count, d_groups, num_hedges, hedges_size, 0,
                        *new_hedges_size, *d_coarse_hedges, *d_coarse_hedges_offsets, *d_coarse_srcs_count);
    hgp::sync(c);
  });
}

hgp_status hgp_coarsen_touching(hgp_ctx* ctx, const uint32_t* d_coarse_hedges, const uint64_t* d_coarse_hedges_offsets,
                                const uint32_t* d_coarse_srcs_count, uint32_t new_num_nodes, uint32_t num_hedges,
                                uint64_t* new_touching_size, uint32_t** d_coarse_touching,
                                uint64_t** d_coarse_touching_offsets, uint32_t** d_coarse_inbound_count) {
  if (!new_touching_size || !d_coarse_touching || !d_coarse_touching_offsets || !d_coarse_inbound_count)
    return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp::build_touching(c, d_coarse_hedges, d_coarse_hedges_offsets, d_coarse_srcs_count, new_num_nodes, num_hedges,
                        *new_touching_size, *d_coarse_touching, *d_coarse_touching_offsets, *d_coarse_inbound_count);
    hgp::sync(c);
  });
}

hgp_status hgp_uncoarsen(hgp_ctx* ctx, const uint32_t* d_groups, const uint32_t* d_coarse_partitions, uint32_t num_nodes,
                         uint32_t* d_partitions) {
  return guard(ctx, [&](Ctx& c) {
    hgp::uncoarsen(c, d_groups, d_coarse_partitions, num_nodes, d_partitions);
    hgp::sync(c);
  });
}

static hgp::RefineLevel make_level(uint32_t n, uint32_t e, uint32_t k, const uint32_t* hedges, const uint64_t* hoff,
                                   const uint32_t* srcc, const uint32_t* tch, const uint64_t* toff, const uint32_t* inb,
                                   const float* w, const uint32_t* sizes) {
  hgp::RefineLevel L{n, e, k, hedges, hoff, srcc, tch, toff, inb, w, sizes};
  return L;
}

hgp_status hgp_refine(hgp_ctx* ctx, const hgp_params* params, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                      const uint32_t* d_srcs_count, const uint32_t* d_touching, const uint64_t* d_touching_offsets,
                      const uint32_t* d_inbound_count, const float* d_hedge_weights, const uint32_t* d_nodes_sizes,
                      uint32_t level_idx, uint32_t num_nodes, uint32_t num_hedges, uint32_t num_partitions,
                      int32_t update_final_inbound_counts, uint32_t* d_partitions, uint32_t* d_partitions_sizes,
                      uint32_t* d_partitions_inbound_sizes) {
  (void)level_idx;
  if (!params || !d_partitions || !d_partitions_sizes || num_partitions == 0) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp::RefineLevel L = make_level(num_nodes, num_hedges, num_partitions, d_hedges, d_hedges_offsets, d_srcs_count,
                                    d_touching, d_touching_offsets, d_inbound_count, d_hedge_weights, d_nodes_sizes);
    hgp::refine(c, L, params->mode, params->max_nodes_per_part, params->max_inbound_per_part, params->refine_repeats,
                update_final_inbound_counts != 0, d_partitions, d_partitions_sizes, d_partitions_inbound_sizes);
    hgp::sync(c);
  });
}

hgp_status hgp_refine_lp(hgp_ctx* ctx, const hgp_params* params, const uint32_t* d_hedges,
                         const uint64_t* d_hedges_offsets, const uint32_t* d_srcs_count, const uint32_t* d_touching,
                         const uint64_t* d_touching_offsets, const uint32_t* d_inbound_count,
                         const float* d_hedge_weights, const uint32_t* d_nodes_sizes, uint32_t num_nodes,
                         uint32_t num_hedges, uint32_t num_partitions, uint32_t* d_partitions,
                         uint32_t* d_partitions_sizes) {
  if (!params || !d_partitions || !d_partitions_sizes || num_partitions == 0) return HGP_E_INVALID;
  if (params->mode != HGP_MODE_KWAY) return HGP_E_INVALID;  // sizes only: no inbound-set checks
  return guard(ctx, [&](Ctx& c) {
    hgp::RefineLevel L = make_level(num_nodes, num_hedges, num_partitions, d_hedges, d_hedges_offsets, d_srcs_count,
                                    d_touching, d_touching_offsets, d_inbound_count, d_hedge_weights, d_nodes_sizes);
    hgp::lp_refine(c, L, params->max_nodes_per_part, params->lp_rounds, d_partitions, d_partitions_sizes);
    hgp::sync(c);
  });
}

hgp_status hgp_refine_analyze(hgp_ctx* ctx, const hgp_params* params, const uint32_t* d_hedges,
                              const uint64_t* d_hedges_offsets, const uint32_t* d_srcs_count, const uint32_t* d_touching,
                              const uint64_t* d_touching_offsets, const uint32_t* d_inbound_count,
                              const float* d_hedge_weights, const uint32_t* d_nodes_sizes, uint32_t num_nodes,
                              uint32_t num_hedges, uint32_t num_partitions, uint32_t fm_repeat, int32_t chainup,
                              const uint32_t* d_partitions, const uint32_t* d_partitions_sizes, uint32_t* h_moves,
                              int64_t* h_gains, uint32_t* h_ranks, int64_t* h_seq_gains, int32_t* h_size_viol,
                              int32_t* h_inbound_viol, uint32_t* h_best_rank) {
  if (!params || num_partitions == 0) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp::RefineLevel L = make_level(num_nodes, num_hedges, num_partitions, d_hedges, d_hedges_offsets, d_srcs_count,
                                    d_touching, d_touching_offsets, d_inbound_count, d_hedge_weights, d_nodes_sizes);
    hgp::AnalyzeOut out;
    out.moves = h_moves;
    out.gains = h_gains;
    out.ranks = h_ranks;
    out.seq = h_seq_gains;
    out.sviol = h_size_viol;
    out.iviol = h_inbound_viol;
    const uint32_t b = hgp::refine_analyze(c, L, params->mode, params->max_nodes_per_part, params->max_inbound_per_part,
                                           params->refine_repeats, fm_repeat, chainup != 0, d_partitions,
                                           d_partitions_sizes, out);
    if (h_best_rank) *h_best_rank = b;
  });
}

hgp_status hgp_greedy_merge_groups(hgp_ctx* ctx, const uint32_t* d_nodes_sizes, const uint32_t* d_inbound_count,
                                   uint32_t num_nodes, uint32_t num_groups, uint32_t max_nodes_per_part,
                                   uint32_t max_inbound_per_part, uint32_t* d_groups, uint32_t* d_groups_sizes,
                                   uint32_t* new_num_groups) {
  if (!new_num_groups) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp::greedy_merge_groups(c, d_nodes_sizes, d_inbound_count, num_nodes, num_groups, max_nodes_per_part,
                             max_inbound_per_part, d_groups, d_groups_sizes, *new_num_groups);
    hgp::sync(c);
  });
}

hgp_status hgp_merge_small_partitions(hgp_ctx* ctx, const uint32_t* d_partitions_sizes,
                                      const uint32_t* d_partitions_inbound_sizes, uint32_t num_nodes,
                                      uint32_t num_partitions, uint32_t max_nodes_per_part, uint32_t max_inbound_per_part,
                                      uint32_t* d_partitions) {
  return guard(ctx, [&](Ctx& c) {
    hgp::merge_small_partitions(c, d_partitions_sizes, d_partitions_inbound_sizes, num_nodes, num_partitions,
                                max_nodes_per_part, max_inbound_per_part, d_partitions);
    hgp::sync(c);
  });
}

hgp_status hgp_relabel_partitions(hgp_ctx* ctx, uint32_t num_nodes, uint32_t* d_partitions, uint32_t* new_num_partitions) {
  if (!new_num_partitions) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp::relabel(c, num_nodes, d_partitions, *new_num_partitions);
    hgp::sync(c);
  });
}

hgp_status hgp_inbound_sizes(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                             const uint32_t* d_srcs_count, const uint32_t* d_partitions, uint32_t num_hedges,
                             uint32_t num_partitions, uint32_t* d_out) {
  return guard(ctx, [&](Ctx& c) {
    hgp::inbound_sizes(c, d_hedges, d_hedges_offsets, d_srcs_count, d_partitions, num_hedges, num_partitions, d_out);
    hgp::sync(c);
  });
}

hgp_status hgp_connectivity(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                            const float* d_hedge_weights, const uint32_t* d_partitions, uint32_t num_hedges,
                            int64_t* km1_fixed) {
  if (!km1_fixed) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    *km1_fixed = hgp::connectivity_fixed(c, d_hedges, d_hedges_offsets, d_hedge_weights, d_partitions, num_hedges);
  });
}

hgp_status hgp_quotient(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                        const uint32_t* d_srcs_count, const float* d_hedge_weights, const uint32_t* d_partitions,
                        uint32_t num_nodes, uint32_t num_hedges, uint32_t* num_partitions, uint32_t* num_merged,
                        uint64_t* num_pins, uint32_t** d_pins, uint64_t** d_offsets, uint32_t** d_srcs_count_out,
                        float** d_weights) {
  if (!num_partitions || !num_merged || !num_pins || !d_pins || !d_offsets || !d_srcs_count_out || !d_weights)
    return HGP_E_INVALID;
  if (num_hedges && (!d_hedges || !d_hedges_offsets || !d_srcs_count || !d_hedge_weights)) return HGP_E_INVALID;
  if (num_nodes && !d_partitions) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    uint32_t k = 0, u = 0;
    uint64_t p = 0;
    hgp::quotient(c, d_hedges, d_hedges_offsets, d_srcs_count, d_hedge_weights, d_partitions, num_nodes, num_hedges, k,
                  u, p, *d_pins, *d_offsets, *d_srcs_count_out, *d_weights);
    hgp::sync(c);
    *num_partitions = k;
    *num_merged = u;
    *num_pins = p;
  });
}

hgp_status hgp_autotune(hgp_ctx* ctx, uint32_t max_log2, hgp_tuning* out) {
  return guard(ctx, [&](Ctx& c) {
    hgp::autotune(c, max_log2);
    if (out) to_public(c.tune, out);
  });
}

hgp_status hgp_tuning_get(const hgp_ctx* ctx, hgp_tuning* out) {
  if (!ctx || !out) return HGP_E_INVALID;
  to_public(ctx->c.tune, out);
  return HGP_OK;
}

hgp_status hgp_tuning_set(hgp_ctx* ctx, const hgp_tuning* in) {
  if (!ctx || !in) return HGP_E_INVALID;
  hgp::Tuning t;
  if (!from_public(in, t)) return HGP_E_INVALID;
  ctx->c.tune = t;
  return HGP_OK;
}

void hgp_tuning_defaults(hgp_tuning* out) {
  if (out) to_public(hgp::Tuning(), out);
}

hgp_status hgp_tuning_set_process_default(const hgp_tuning* in) {
  hgp::Tuning t;
  if (in && !from_public(in, t)) return HGP_E_INVALID;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_default_tuning = t;
  return HGP_OK;
}

hgp_status hgp_tuning_save(const hgp_tuning* t, const char* path) {
  if (!t || !path) return HGP_E_INVALID;
  std::ofstream f(path);
  if (!f) return HGP_E_IO;
  f << "# hgp_tuning (launch parameters; results do not depend on them)\n";
  for (int i = 0; i < 40; ++i)
    if (t->sort_rounds[i]) f << "sort_rounds " << i << " " << t->sort_rounds[i] << "\n";
  f << "heavy_table_slots " << t->heavy_table_slots << "\n";
  f << "heavy_table_fill16 " << t->heavy_table_fill16 << "\n";
  f << "cascade_block " << t->cascade_block << "\n";
  f << "scan_lookback_tiles " << t->scan_lookback_tiles << "\n";
  f << "mask_wide4_degree " << t->mask_wide4_degree << "\n";
  f << "mask_wide16_degree " << t->mask_wide16_degree << "\n";
  f << "side_streams " << t->side_streams << "\n";
  f << "coop_max_n " << t->coop_max_n << "\n";
  f << "coop_tile " << t->coop_tile << "\n";
  f << "tuned " << t->tuned << "\n";
  return f.good() ? HGP_OK : HGP_E_IO;
}

hgp_status hgp_tuning_load(const char* path, hgp_tuning* out) {
  if (!path || !out) return HGP_E_INVALID;
  std::ifstream f(path);
  if (!f) return HGP_E_IO;
  hgp_tuning t;
  hgp_tuning_defaults(&t);
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    std::string key;
    ss >> key;
    uint64_t a = 0, b = 0;
    if (key == "sort_rounds") {
      if (!(ss >> a >> b) || a >= 40 || b > 16) return HGP_E_INVALID;
      t.sort_rounds[a] = (uint32_t)b;
      continue;
    }
    if (!(ss >> a) || a > 0xFFFFFFFFull) return HGP_E_INVALID;
    if (key == "heavy_table_slots") t.heavy_table_slots = (uint32_t)a;
    else if (key == "heavy_table_fill16") t.heavy_table_fill16 = (uint32_t)a;
    else if (key == "cascade_block") t.cascade_block = (uint32_t)a;
    else if (key == "scan_lookback_tiles") t.scan_lookback_tiles = (uint32_t)a;
    else if (key == "mask_wide4_degree") t.mask_wide4_degree = (uint32_t)a;
    else if (key == "mask_wide16_degree") t.mask_wide16_degree = (uint32_t)a;
    else if (key == "side_streams") t.side_streams = (uint32_t)a;
    else if (key == "coop_max_n") t.coop_max_n = (uint32_t)a;
    else if (key == "coop_tile") t.coop_tile = (uint32_t)a;
    else if (key == "tuned") t.tuned = (uint32_t)a;
    else return HGP_E_INVALID;
  }
  hgp::Tuning chk;
  if (!from_public(&t, chk)) return HGP_E_INVALID;
  *out = t;
  return HGP_OK;
}

hgp_status hgp_connectivity_noalloc(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                                    const float* d_hedge_weights, const uint32_t* d_partitions, uint32_t num_hedges,
                                    uint32_t num_partitions, int64_t* d_km1_fixed) {
  if (!d_km1_fixed || (num_hedges && (!d_hedges || !d_hedges_offsets || !d_hedge_weights || !d_partitions)))
    return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp::connectivity_noalloc(c, d_hedges, d_hedges_offsets, d_hedge_weights, d_partitions, num_hedges, num_partitions,
                              d_km1_fixed);
  });
}

hgp_status hgp_initial_partition(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                                 const float* d_hedge_weights, uint32_t num_nodes, uint32_t num_hedges,
                                 uint64_t hedges_size, const uint32_t* d_nodes_sizes, uint32_t k, float epsilon,
                                 uint32_t max_nodes_per_part, int32_t prune, uint32_t** d_partitions,
                                 uint32_t** d_partitions_sizes) {
  (void)hedges_size;
  if (!d_partitions || !d_partitions_sizes || k == 0) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp::initial_partition(c, d_hedges, d_hedges_offsets, d_hedge_weights, num_nodes, num_hedges, d_nodes_sizes, k,
                           epsilon, max_nodes_per_part, prune, *d_partitions, *d_partitions_sizes, nullptr);
  });
}

hgp_status hgp_ip_input(hgp_ctx* ctx, const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                        const float* d_hedge_weights, const uint32_t* d_nodes_sizes, uint32_t num_nodes,
                        uint32_t num_hedges, int32_t prune, uint32_t* num_merged, uint64_t* merged_pins,
                        uint64_t* h_offsets, uint32_t* h_pins, uint64_t* h_weights) {
  if (!num_merged || !merged_pins) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    hgp_ip::Hypergraph h;
    hgp::ip_input_device(c, d_hedges, d_hedges_offsets, d_hedge_weights, d_nodes_sizes, num_nodes, num_hedges,
                         (uint32_t)std::max(prune, 0), h);
    *num_merged = (uint32_t)(h.off.size() - 1);
    *merged_pins = h.pins.size();
    if (h_offsets) std::copy(h.off.begin(), h.off.end(), h_offsets);
    if (h_pins) std::copy(h.pins.begin(), h.pins.end(), h_pins);
    if (h_weights) std::copy(h.w.begin(), h.w.end(), h_weights);
  });
}

hgp_status hgp_partition(hgp_ctx* ctx, const hgp_params* params, const hgp_hgraph* hg, uint32_t* h_partitions_out,
                         hgp_run_stats* stats) {
  if (!params || !hg || !h_partitions_out) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    const int saved = c.verbose;
    c.verbose = params->verbose;
    try {
      hgp::PipelineInput in;
      in.n = hg->n;
      in.e = hg->e();
      in.p = hg->pins.size();
      in.host = hg;
      hgp::partition(c, *params, in, h_partitions_out, stats);
    } catch (...) {
      c.verbose = saved;
      throw;
    }
    c.verbose = saved;
  });
}

hgp_status hgp_profile_kernel(hgp_ctx* ctx, const char* kernel_name) {
  return guard(ctx, [&](Ctx& c) {
    c.prof_kernel = kernel_name ? kernel_name : "";
    if (!c.prof_kernel.empty() && !c.prof_dev) {
      c.prof_dev = (unsigned long long*)hgp::dmalloc(c, hgp::PROF_SLOTS * 8);
      hgp::memset_async(c, c.prof_dev, 0, hgp::PROF_SLOTS * 8);
    }
  });
}

hgp_status hgp_profile_count(hgp_ctx* ctx, int32_t on) { return guard(ctx, [&](Ctx& c) { c.prof_count = on != 0; }); }

hgp_status hgp_profile_read(hgp_ctx* ctx, double* total_ms, uint64_t* launches, uint64_t* algorithmic_bytes) {
  return guard(ctx, [&](Ctx& c) {
    hgp::sync(c);
    double ms = 0.0;
    uint64_t bytes = 0;
    for (auto& r : c.prof) {
      float t = 0.0f;
      HGP_HIP(hipEventElapsedTime(&t, r.a, r.b));
      ms += t;
      bytes += r.bytes;
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    if (c.prof_dev) {  // bytes the kernel counted on the device
      unsigned long long slots[hgp::PROF_SLOTS];
      hgp::copy_d2h(c, slots, c.prof_dev, sizeof(slots));
      for (auto v : slots) bytes += v;
      hgp::memset_async(c, c.prof_dev, 0, sizeof(slots));
      hgp::sync(c);
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = c.prof.size();
    if (algorithmic_bytes) *algorithmic_bytes = bytes;
    c.prof.clear();
  });
}

hgp_status hgp_partition_device(hgp_ctx* ctx, const hgp_params* params, uint32_t num_nodes, uint32_t num_hedges,
                                const uint32_t* d_hedges, const uint64_t* d_hedges_offsets,
                                const uint32_t* d_srcs_count, const float* d_hedge_weights,
                                uint32_t* h_partitions_out, hgp_run_stats* stats) {
  if (!params || !h_partitions_out || !d_hedges_offsets) return HGP_E_INVALID;
  return guard(ctx, [&](Ctx& c) {
    const int saved = c.verbose;
    c.verbose = params->verbose;
    try {
      hgp::PipelineInput in;
      in.n = num_nodes;
      in.e = num_hedges;
      // the offsets start at 0 and the pins name nodes below num_nodes (the host path
      // checks the same when it builds its model)
      HGP_REQUIRE(hgp::read_scalar(c, d_hedges_offsets) == 0, HGP_E_INVALID,
                  "hgp_partition_device: d_hedges_offsets[0] must be 0");
      in.p = num_hedges ? hgp::read_scalar(c, d_hedges_offsets + num_hedges) : 0;
      HGP_REQUIRE(in.p == 0 || d_hedges, HGP_E_INVALID, "hgp_partition_device: pins without a pin array");
      HGP_REQUIRE(in.p == 0 || hgp::reduce_max_u32(c, d_hedges, in.p) < num_nodes, HGP_E_INVALID,
                  "hgp_partition_device: a pin is not below num_nodes");
      in.d_hedges = d_hedges;
      in.d_hoff = d_hedges_offsets;
      in.d_srcc = d_srcs_count;
      in.d_w = d_hedge_weights;
      hgp::partition(c, *params, in, h_partitions_out, stats);
    } catch (...) {
      c.verbose = saved;
      throw;
    }
    c.verbose = saved;
  });
}

}  // extern "C"
