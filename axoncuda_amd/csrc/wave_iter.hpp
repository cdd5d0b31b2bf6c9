// wave_iter.hpp - wave-cooperative traversal of "all pins of all hedges touching a node".
//
// The hedges of a touching list have very different sizes (2 .. 1000+ pins); giving one
// hedge to one lane or one hedge to the whole wave wastes most lanes.  Instead the wave
// takes 64 touching entries at a time, prefix-sums their hedge sizes across lanes, and
// then walks the flattened pin range 64 pins per step; each lane finds the hedge that
// owns its pin with a 6-step shuffle binary search.  Every lane calls the visitor on every
// step (with `act` false past the end) so visitors may use wave-wide ballots.
#pragma once

#include "common.hpp"

namespace hgp {

struct PinVisit {
  bool act;         // lane holds a real pin
  uint64_t j;       // index into the touching list
  uint32_t e;       // hedge id
  uint32_t pos;     // position of the pin inside the hedge
  uint32_t len;     // hedge size
  uint32_t pin;     // node id
};

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, uint32_t src) {
  uint32_t lo = __shfl((uint32_t)v, (int)src, 64);
  uint32_t hi = __shfl((uint32_t)(v >> 32), (int)src, 64);
  return ((uint64_t)hi << 32) | lo;
}

// Owner lane of flat position f0 + lane among the 64 hedges of a chunk whose exclusive start
// offsets are `excl` (each hedge has at least one entry, so starts are distinct; lanes past the
// list carry length 0 and start at the chunk total, beyond every active position): the number
// of starts before the window (a ballot) plus those inside it at or before this position,
// minus one.  The starts inside the window belong to lanes [before, upto), a run because the
// starts ascend with the lane.  A window of long hedges holds few of them, and each lane
// compares its position with each one (read as a scalar); a window of many short hedges
// counts them from an OR-reduced bitmap of the window's starts.  (Positions past the total
// may get another owner lane than the bitmap would give; they are inactive.)
constexpr uint32_t OWNER_FEW_STARTS = 12;
__device__ __forceinline__ uint32_t chunk_owner(uint32_t excl, uint32_t f0, uint64_t le_l) {
  const uint32_t before = (uint32_t)__popcll(__ballot(excl < f0));
  const uint32_t upto = (uint32_t)__popcll(__ballot(excl < f0 + 64u));
  if (upto - before <= OWNER_FEW_STARTS) {  // uniform over the wave
    const uint32_t fi = f0 + lane_id();
    uint32_t k = before - 1u;  // before >= 1 unless f0 == 0, where lane 0 starts at 0 <= fi
    for (uint32_t x = before; x < upto; ++x) k += (uint32_t)__builtin_amdgcn_readlane((int)excl, (int)x) <= fi ? 1u : 0u;
    return k;
  }
  const uint64_t mine = (excl >= f0 && excl - f0 < 64u) ? (1ull << (excl - f0)) : 0ull;
  const uint64_t heads = wave_reduce(mine, [](uint64_t a, uint64_t b) { return a | b; });
  return before + (uint32_t)__popcll(heads & le_l) - 1u;
}
__device__ __forceinline__ uint64_t lanes_upto(uint32_t l) { return l == 63 ? ~0ull : ((2ull << l) - 1ull); }

// Generic form: hedge e owns the entry range [start(e), start(e) + length(e)) of some
// per-hedge array; the visitor gets the flattened entry index in `pin` replaced by the
// absolute entry index (field `abs`).
struct EntryVisit {
  bool act;
  uint64_t j;    // index into the touching list
  uint32_t e;    // hedge id
  uint32_t pos;  // position inside the hedge's range
  uint64_t abs;  // absolute entry index (wave_visit_entries only)
  uint32_t w;    // the hedge's fixed-point weight (wave_visit_hedge_lists only)
  uint2 pe;      // the list entry itself (wave_visit_hedge_lists only)
};

template <typename StartF, typename LenF, typename F>
__device__ __forceinline__ void wave_visit_entries(const uint32_t* __restrict__ tch, uint64_t t0, uint64_t t1,
                                                   StartF&& start_of, LenF&& len_of, F&& visit) {
  const uint32_t l = lane_id();
  for (uint64_t jb = t0; jb < t1; jb += 64) {
    const uint64_t j = jb + l;
    uint32_t e = 0, len = 0;
    uint64_t s = 0;
    if (j < t1) {
      e = tch[j];
      s = start_of(e);
      len = len_of(e);
    }
    const uint32_t incl = wave_scan_incl(len);
    const uint32_t total = __shfl(incl, 63, 64);
    const uint32_t excl = incl - len;
    for (uint32_t f0 = 0; f0 < total; f0 += 64) {
      const uint32_t fi = f0 + l;
      uint32_t k = 0;
#pragma unroll
      for (uint32_t step = 32; step > 0; step >>= 1) {
        const uint32_t cand = k + step;
        const uint32_t ex = __shfl(excl, (int)cand, 64);
        if (ex <= fi) k = cand;
      }
      EntryVisit ev;
      ev.act = fi < total;
      ev.e = __shfl(e, (int)k, 64);
      const uint64_t os = shfl_u64(s, k);
      const uint32_t oex = __shfl(excl, (int)k, 64);
      ev.j = jb + k;
      ev.pos = fi - oex;
      ev.abs = os + ev.pos;
      visit(ev);
    }
  }
}

// The same traversal over per-hedge lists of 8-byte entries described by one 16-byte record
// per hedge (start lo, start hi, length, fixed-point weight), so each touching hedge costs one
// load.  A list of one entry is held in the record itself, in place of the start, and is not
// read from `lists`.  The visitor gets the entry in `pe`.  Returns this lane's share of the
// summed weights of the touching hedges.
template <typename F>
__device__ __forceinline__ int64_t wave_visit_hedge_lists(const uint32_t* __restrict__ tch, uint64_t t0, uint64_t t1,
                                                          const uint4* __restrict__ hinfo,
                                                          const uint2* __restrict__ lists, F&& visit) {
  const uint32_t l = lane_id();
  int64_t wsum = 0;
  for (uint64_t jb = t0; jb < t1; jb += 64) {
    const uint64_t j = jb + l;
    uint32_t e = 0, len = 0, wf = 0;
    uint64_t s = 0;
    if (j < t1) {
      e = tch[j];
      const uint4 hi = hinfo[e];
      s = ((uint64_t)hi.y << 32) | hi.x;
      len = hi.z;
      wf = hi.w;
      wsum += wf;
    }
    const uint32_t incl = wave_scan_incl(len);
    const uint32_t total = __shfl(incl, 63, 64);
    const uint32_t excl = incl - len;
    const uint64_t le_l = lanes_upto(l);
    for (uint32_t f0 = 0; f0 < total; f0 += 64) {
      const uint32_t fi = f0 + l;
      const uint32_t k = chunk_owner(excl, f0, le_l);  // every touching hedge lists >= 1 entry
      EntryVisit ev;
      ev.act = fi < total;
      ev.e = __shfl(e, (int)k, 64);
      ev.w = __shfl(wf, (int)k, 64);
      const uint64_t os = shfl_u64(s, k);
      const uint32_t oex = __shfl(excl, (int)k, 64);
      const uint32_t olen = __shfl(len, (int)k, 64);
      ev.j = jb + k;
      ev.pos = fi - oex;
      ev.abs = 0;
      ev.pe = olen == 1 ? make_uint2((uint32_t)os, (uint32_t)(os >> 32))
                        : (ev.act ? lists[os + ev.pos] : make_uint2(0u, 0u));
      visit(ev);
    }
  }
  return wsum;
}

// `stop()` (wave-uniform) is checked before every step; the traversal ends when it is true.
template <typename F, typename S>
__device__ __forceinline__ void wave_visit_pins_until(const uint32_t* __restrict__ tch, uint64_t t0, uint64_t t1,
                                                      const uint64_t* __restrict__ hoff,
                                                      const uint32_t* __restrict__ hedges, F&& visit, S&& stop) {
  const uint32_t l = lane_id();
  for (uint64_t jb = t0; jb < t1; jb += 64) {
    if (stop()) return;
    const uint64_t j = jb + l;
    uint32_t e = 0, len = 0;
    uint64_t s = 0;
    if (j < t1) {
      e = tch[j];
      s = hoff[e];
      len = (uint32_t)(hoff[e + 1] - s);
    }
    const uint32_t incl = wave_scan_incl(len);
    const uint32_t total = __shfl(incl, 63, 64);
    const uint32_t excl = incl - len;
    const uint64_t le_l = lanes_upto(l);
    for (uint32_t f0 = 0; f0 < total; f0 += 64) {
      if (f0 && stop()) return;
      const uint32_t fi = f0 + l;
      const uint32_t k = chunk_owner(excl, f0, le_l);  // every touching hedge has a pin
      PinVisit pv;
      pv.act = fi < total;
      pv.e = __shfl(e, (int)k, 64);
      const uint64_t os = shfl_u64(s, k);
      const uint32_t oex = __shfl(excl, (int)k, 64);
      pv.len = __shfl(len, (int)k, 64);
      pv.j = jb + k;
      pv.pos = fi - oex;
      pv.pin = pv.act ? hedges[os + pv.pos] : 0u;
      visit(pv);
    }
  }
}

// The same traversal with two per-hedge values computed once per touching hedge by the lane
// that loads it, `hvf(e, len)`, and handed to every visit of that hedge's pins in `hv`.  With
// PIPE the next step's pin is loading while the current step is visited (one more step's
// registers live across the visit).
struct PinVisitHV : PinVisit {
  uint2 hv;
};
template <bool PIPE = true, typename H, typename F, typename S>
__device__ __forceinline__ void wave_visit_pins_hv_until(const uint32_t* __restrict__ tch, uint64_t t0, uint64_t t1,
                                                         const uint64_t* __restrict__ hoff,
                                                         const uint32_t* __restrict__ hedges, H&& hvf, F&& visit,
                                                         S&& stop) {
  const uint32_t l = lane_id();
  for (uint64_t jb = t0; jb < t1; jb += 64) {
    if (stop()) return;
    const uint64_t j = jb + l;
    uint32_t e = 0, len = 0;
    uint64_t s = 0;
    uint2 hv = make_uint2(0u, 0u);
    if (j < t1) {
      e = tch[j];
      s = hoff[e];
      len = (uint32_t)(hoff[e + 1] - s);
      hv = hvf(e, len);
    }
    const uint32_t incl = wave_scan_incl(len);
    const uint32_t total = __shfl(incl, 63, 64);
    const uint32_t excl = incl - len;
    const uint64_t le_l = lanes_upto(l);
    // step f0's owner, fields and pin load
    auto step = [&](uint32_t f0) {
      const uint32_t fi = f0 + l;
      const uint32_t k = chunk_owner(excl, f0, le_l);
      PinVisitHV pv;
      pv.act = fi < total;
      pv.e = __shfl(e, (int)k, 64);
      const uint64_t os = shfl_u64(s, k);
      const uint32_t oex = __shfl(excl, (int)k, 64);
      pv.len = __shfl(len, (int)k, 64);
      pv.hv.x = __shfl(hv.x, (int)k, 64);
      pv.hv.y = __shfl(hv.y, (int)k, 64);
      pv.j = jb + k;
      pv.pos = fi - oex;
      pv.pin = pv.act ? hedges[os + pv.pos] : 0u;
      return pv;
    };
    if constexpr (PIPE) {
      PinVisitHV cur = step(0);
      for (uint32_t f0 = 0; f0 < total; f0 += 64) {
        if (f0 && stop()) return;
        PinVisitHV nxt = cur;
        if (f0 + 64 < total) nxt = step(f0 + 64);  // uniform: total is wave-wide
        visit(cur);
        cur = nxt;
      }
    } else {
      for (uint32_t f0 = 0; f0 < total; f0 += 64) {
        if (f0 && stop()) return;
        visit(step(f0));
      }
    }
  }
}

template <typename F>
__device__ __forceinline__ void wave_visit_pins(const uint32_t* __restrict__ tch, uint64_t t0, uint64_t t1,
                                                const uint64_t* __restrict__ hoff,
                                                const uint32_t* __restrict__ hedges, F&& visit) {
  wave_visit_pins_until(tch, t0, t1, hoff, hedges, visit, [] { return false; });
}

}  // namespace hgp
