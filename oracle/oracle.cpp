// oracle.cpp - sequential CPU restatement of the reference hgraph_gpu pipeline.
//
// TEST INFRASTRUCTURE ONLY: the parity checker for the HIP implementation.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library;
// the product never does.  Every function restates one reference step on host arrays
// (file:line relative to the reference repository cited at each definition) and emits
// the canonical orders of include/hgp.h.  Where the reference is timing-dependent
// (cooperative pairing) or lane-dependent (gain ties, float sums), the documented
// deterministic rule that the HIP kernels also implement is restated instead.
//
// Pinning (what checks each function; see tests/):
//   loaders / dedup / metrics / validity : against the reference's own host C++
//     (includes/hgraph.hpp, sources/constr.cpp, grading/grade.cpp) built into oracle/_ref,
//     and the committed golden fixtures produced from it;
//   touching / neighbours / coarse hedges : brute-force set definitions on tiny inputs
//     (transpose, union of pins, set of groups) and invariants (sizes, sums);
//   candidates : brute-force pairwise score enumeration + symmetry invariant;
//   pairing    : exhaustive maximum-weight matching on tiny forests;
//   refinement : direct sequential simulation of each prefix (recount lambda, sizes,
//                inbound sets) on tiny inputs;
//   merges / relabel : closed-form greedy restatement + invariants.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "ip_core.hpp"

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;

// ---------------------------------------------------------------------------------------
// host model
struct HG {
  uint32_t n = 0;
  std::vector<uint64_t> off{0};
  std::vector<uint32_t> pins;
  std::vector<uint32_t> srcc;
  std::vector<float> w;
  uint32_t e() const { return (uint32_t)srcc.size(); }
};

bool ends_with(const std::string& s, const char* suf) {
  size_t l = strlen(suf);
  return s.size() >= l && s.compare(s.size() - l, l, suf) == 0;
}

// loadhMETIS (includes/hgraph.hpp:936-1013): first non-comment line "E V [fmt]",
// 1-based pins, fmt 1/11 carry a leading hedge weight, vertex weights ignored,
// single-pin hedges skipped; all pins are destinations (src_count 0).
bool load_hgr(const std::string& path, HG& hg) {
  std::ifstream f(path);
  if (!f) return false;
  auto next_line = [&](std::string& out) -> bool {
    std::string line;
    while (std::getline(f, line)) {
      if (!line.empty() && line.back() == '\r') line.pop_back();
      size_t i = line.find_first_not_of(" \t");
      if (i == std::string::npos) continue;
      if (line[i] == '%') continue;
      out = line.substr(i);
      return true;
    }
    return false;
  };
  std::string header;
  if (!next_line(header)) return false;
  std::istringstream hs(header);
  uint64_t E = 0, V = 0;
  uint32_t fmt = 0;
  hs >> E >> V;
  if (E > 0xFFFFFFFFull || V > 0xFFFFFFFFull) return false;
  if (!(hs >> fmt)) fmt = 0;
  bool ew = (fmt == 1 || fmt == 11);
  hg.n = (uint32_t)V;
  for (uint64_t i = 0; i < E; ++i) {
    std::string line;
    if (!next_line(line)) return false;
    std::istringstream ls(line);
    float w = 1.0f;
    if (ew) ls >> w;
    std::vector<uint32_t> nodes;
    uint32_t v;
    while (ls >> v) {
      if (v == 0 || v > V) return false;
      nodes.push_back(v - 1);
    }
    if (nodes.empty()) return false;
    if (nodes.size() == 1) continue;
    hg.pins.insert(hg.pins.end(), nodes.begin(), nodes.end());
    hg.off.push_back(hg.pins.size());
    hg.srcc.push_back(0);
    hg.w.push_back(w);
  }
  return true;
}

// loadSNN (:707-743): u32 N, E; per hedge u32 dstc, u32 src, u32 dst[dstc], f32 w.
// loadAXH (:805-854): u32 N, E; per hedge u32 pins, u32 srcs, u32 pin[pins], f32 w.
bool load_bin(const std::string& path, bool axh, HG& hg) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  auto r32 = [&](uint32_t& v) { return (bool)f.read((char*)&v, 4); };
  uint32_t N, E;
  if (!r32(N) || !r32(E)) return false;
  hg.n = N;
  for (uint32_t i = 0; i < E; ++i) {
    uint32_t a, b;
    if (!r32(a) || !r32(b)) return false;
    uint32_t npins, nsrc;
    if (axh) {
      npins = a;
      nsrc = b;
      if (nsrc == 0 || npins < nsrc) return false;
      for (uint32_t k = 0; k < npins; ++k) {
        uint32_t v;
        if (!r32(v) || v >= N) return false;
        hg.pins.push_back(v);
      }
    } else {
      npins = a + 1;
      nsrc = 1;
      if (b >= N) return false;
      hg.pins.push_back(b);
      for (uint32_t k = 0; k < a; ++k) {
        uint32_t v;
        if (!r32(v) || v >= N) return false;
        hg.pins.push_back(v);
      }
    }
    if (npins - nsrc < 1) return false;  // HyperGraph(): at least one destination
    float w;
    if (!f.read((char*)&w, 4)) return false;
    hg.off.push_back(hg.pins.size());
    hg.srcc.push_back(nsrc);
    hg.w.push_back(w);
  }
  return true;
}

// deduplicateHyperedges(2,false) (includes/hgraph.hpp:518-581): sources first, each pin
// kept at its first occurrence, a destination equal to a source dropped.
uint64_t dedup_rule2(uint32_t E, uint64_t* off, uint32_t* pins, uint32_t* srcc) {
  uint64_t w = 0;
  uint64_t prev_end = off[0];
  std::unordered_set<uint32_t> seen;
  for (uint32_t i = 0; i < E; ++i) {
    uint64_t s = prev_end, t = off[i + 1];
    prev_end = t;
    uint64_t start = w;
    seen.clear();
    uint32_t sc = srcc[i];
    for (uint64_t k = s; k < s + sc; ++k)
      if (seen.insert(pins[k]).second) pins[w++] = pins[k];
    uint32_t nsrc = (uint32_t)(w - start);
    for (uint64_t k = s + sc; k < t; ++k)
      if (seen.insert(pins[k]).second) pins[w++] = pins[k];
    off[i] = start;
    srcc[i] = nsrc;
  }
  off[E] = w;
  return w;
}

inline uint32_t fixed_weight(float w) {
  float v = w * 262144.0f;
  if (!(v > 0.0f)) return 0u;
  if (v >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)v;
}

// Symmetric pairing noise in [0, 64): headers/utils.cuh:159-171 (same mixing steps).
inline uint32_t noise64(uint32_t a, uint32_t b) {
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

// Integer mixer used for gain tie-breaks: headers/utils.cuh:252-261 (same steps).
inline uint32_t mix32(uint32_t x) {
  x ^= x >> 17;
  x *= 0xED5AD4BBu;
  x ^= x >> 11;
  x *= 0xAC4C1B51u;
  x ^= x >> 15;
  x *= 0x31848BABu;
  x ^= x >> 14;
  return x;
}

// ---------------------------------------------------------------------------------------
// construction

// buildTouching (sources/construction.cu:64-179): inbound = hedges where v is a
// destination, outbound = hedges where v is a source; canonical order both ascending.
void touching(uint32_t n, uint32_t E, const uint64_t* off, const uint32_t* pins, const uint32_t* srcc,
              std::vector<uint64_t>& toff, std::vector<uint32_t>& tch, std::vector<uint32_t>& inb) {
  std::vector<uint32_t> in_cnt(n, 0), out_cnt(n, 0);
  for (uint32_t e = 0; e < E; ++e)
    for (uint64_t i = off[e]; i < off[e + 1]; ++i) {
      if (i - off[e] < srcc[e]) out_cnt[pins[i]]++;
      else in_cnt[pins[i]]++;
    }
  toff.assign(n + 1, 0);
  for (uint32_t v = 0; v < n; ++v) toff[v + 1] = toff[v] + in_cnt[v] + out_cnt[v];
  tch.assign(toff[n], 0);
  inb = in_cnt;
  std::vector<uint64_t> pin_in(n), pin_out(n);
  for (uint32_t v = 0; v < n; ++v) {
    pin_in[v] = toff[v];
    pin_out[v] = toff[v] + in_cnt[v];
  }
  for (uint32_t e = 0; e < E; ++e)  // increasing e -> ascending order per node
    for (uint64_t i = off[e]; i < off[e + 1]; ++i) {
      uint32_t v = pins[i];
      if (i - off[e] < srcc[e]) tch[pin_out[v]++] = e;
      else tch[pin_in[v]++] = e;
    }
}

// ---------------------------------------------------------------------------------------
// pairing (groupNodes, sources/coarsening.cu:70-210)
//
// The reference walks each round's pointer forest with racing atomics
// (kernels/coarsening.cu:241-505).  Restated deterministically: in round i every
// unlocked node u points at t = pairs[u][i] when t is also unlocked; cycles are broken
// by making their minimum id a root; then an exact maximum-weight matching of the forest
// is computed bottom-up (with(v) = w(v) + S(v), wout(v) = S(v) + max(0, best child
// gain), S(v) = sum of children's wout, child gain = with - wout, best child by
// (gain, id) descending) and decided top-down: a node not matched to its parent takes
// its best child when that gain is positive.  A pair takes the child's id as raw group
// id (the id written by the reference's lock step, kernels/coarsening.cu:457-462).
void pairing(uint32_t n, const uint32_t* pairs, const uint32_t* scores, uint32_t cnc, std::vector<uint32_t>& raw) {
  raw.assign(n, NONE);
  std::vector<uint8_t> locked(n, 0);
  std::vector<uint32_t> parent(n), w(n), holder(n), pending(n);
  std::vector<uint64_t> S(n), best(n);
  std::vector<uint8_t> done(n);
  std::vector<std::vector<uint32_t>> children(n);
  for (uint32_t i = 0; i < cnc; ++i) {
    for (uint32_t v = 0; v < n; ++v) {
      parent[v] = NONE;
      w[v] = 0;
      if (locked[v]) continue;
      uint32_t t = pairs[(uint64_t)v * cnc + i];
      if (t != NONE && !locked[t]) {
        parent[v] = t;
        w[v] = scores[(uint64_t)v * cnc + i];
      }
    }
    // break cycles: nodes on a cycle of the functional graph; min id becomes a root
    {
      std::vector<uint8_t> color(n, 0);  // 0 new, 1 on stack, 2 finished
      for (uint32_t s = 0; s < n; ++s) {
        if (color[s] || locked[s]) continue;
        std::vector<uint32_t> path;
        uint32_t v = s;
        while (v != NONE && color[v] == 0) {
          color[v] = 1;
          path.push_back(v);
          v = parent[v];
        }
        if (v != NONE && color[v] == 1) {  // found a cycle starting at v
          uint32_t m = v;
          uint32_t x = parent[v];
          while (x != v) {
            m = std::min(m, x);
            x = parent[x];
          }
          parent[m] = NONE;
        }
        for (uint32_t p : path) color[p] = 2;
      }
    }
    // bottom-up DP in child-before-parent order (Kahn on child counts)
    for (uint32_t v = 0; v < n; ++v) {
      children[v].clear();
      S[v] = 0;
      best[v] = 0;
      done[v] = 0;
      pending[v] = 0;
    }
    for (uint32_t v = 0; v < n; ++v)
      if (!locked[v] && parent[v] != NONE) {
        children[parent[v]].push_back(v);
        pending[parent[v]]++;
      }
    std::vector<uint32_t> stack;
    for (uint32_t v = 0; v < n; ++v)
      if (!locked[v] && pending[v] == 0) stack.push_back(v);
    while (!stack.empty()) {
      uint32_t v = stack.back();
      stack.pop_back();
      done[v] = 1;
      uint64_t bg = best[v] >> 32;
      uint64_t wout = S[v] + bg;
      uint32_t p = parent[v];
      if (p == NONE) continue;
      int64_t gain = (int64_t)w[v] - (int64_t)bg;
      uint64_t key = ((uint64_t)(gain > 0 ? gain : 0) << 32) | v;
      S[p] += wout;
      best[p] = std::max(best[p], key);
      if (--pending[p] == 0) stack.push_back(p);
    }
    for (uint32_t v = 0; v < n; ++v)
      holder[v] = (!locked[v] && (best[v] >> 32) > 0) ? (uint32_t)(best[v] & 0xFFFFFFFFu) : NONE;
    // top-down decisions along holder chains
    for (uint32_t y = 0; y < n; ++y) {
      if (locked[y]) continue;
      bool top = parent[y] == NONE || holder[parent[y]] != y;
      if (!top) continue;
      uint32_t cur = y;
      while (cur != NONE) {
        uint32_t h = holder[cur];
        if (h == NONE) break;
        raw[cur] = h;
        raw[h] = h;
        cur = holder[h];
      }
    }
    for (uint32_t v = 0; v < n; ++v)
      if (raw[v] != NONE) locked[v] = 1;
  }
  for (uint32_t v = 0; v < n; ++v)
    if (raw[v] == NONE) raw[v] = v;
}

// build_orphan_pairs (sources/chaining.cu:260-329; kernels/chaining.cu:237-270):
// candidate-less nodes sorted by (size, id), k-th smallest paired with k-th largest
// when the summed sizes and summed inbound counts fit; group id = smaller id.
void orphans(uint32_t n, const uint32_t* pairs, uint32_t cnc, const uint32_t* sizes, const uint32_t* inb,
             uint32_t omega, uint32_t delta, std::vector<uint32_t>& raw) {
  std::vector<uint64_t> keys;
  for (uint32_t v = 0; v < n; ++v)
    if (pairs[(uint64_t)v * cnc] == NONE) keys.push_back(((uint64_t)sizes[v] << 32) | v);
  if (keys.size() < 2) return;
  std::sort(keys.begin(), keys.end());
  size_t m = keys.size();
  for (size_t k = 0; k < m / 2; ++k) {
    uint32_t L = (uint32_t)keys[k], R = (uint32_t)keys[m - 1 - k];
    if ((uint32_t)(sizes[L] + sizes[R]) > omega) continue;
    if ((uint32_t)(inb[L] + inb[R]) > delta) continue;
    uint32_t g = L < R ? L : R;
    raw[L] = g;
    raw[R] = g;
  }
}

// ---------------------------------------------------------------------------------------
// refinement (refinementRepeats, sources/refinement.cu:18-582)

struct Level {
  uint32_t n, E;
  const uint64_t* hoff;
  const uint32_t* pins;
  const uint32_t* srcc;
  const uint64_t* toff;
  const uint32_t* tch;
  const uint32_t* inb;
  const uint32_t* sizes;
  std::vector<uint32_t> W;  // fixed-point weights
};

struct Analysis {
  std::vector<uint32_t> moves, ranks;
  std::vector<int64_t> gains, seq;
  std::vector<int32_t> sviol, iviol;
  uint32_t V = 0;
  uint32_t best = NONE;
};

inline uint32_t discount_of(uint32_t fm_repeat, uint32_t R) {
  return fm_repeat < R / 3 ? 1u : (fm_repeat < 2 * R / 3 ? 2u : NONE);
}

// Per-hedge partition counts: all pins and destination pins.
void hedge_counts(const Level& L, const uint32_t* part, std::vector<std::unordered_map<uint32_t, uint32_t>>& cnt,
                  std::vector<std::unordered_map<uint32_t, uint32_t>>* dcnt) {
  cnt.assign(L.E, {});
  if (dcnt) dcnt->assign(L.E, {});
  for (uint32_t e = 0; e < L.E; ++e)
    for (uint64_t i = L.hoff[e]; i < L.hoff[e + 1]; ++i) {
      uint32_t q = part[L.pins[i]];
      cnt[e][q]++;
      if (dcnt && i - L.hoff[e] >= L.srcc[e]) (*dcnt)[e][q]++;
    }
}

// chaining (sources/chaining.cu:19-254; kernels/chaining.cu:31-230) with integer keys:
// successor score (gain, -|size difference|) replaces w - alpha*|ds|; claims are won by
// (score, -predecessor); components are ranked by (summed gain desc, head id asc) and
// moves ordered by (component rank, position, sorted index).
struct ChainLinks {  // per node: predecessor node (NONE = head), component head node, position
  std::vector<uint32_t> prev, comp, pos;
};

void chain_ranks(uint32_t n, const uint32_t* part, const std::vector<uint32_t>& moves,
                 const std::vector<int64_t>& gains, const uint32_t* sizes, std::vector<uint32_t>& ranks,
                 ChainLinks* links = nullptr) {
  // sorted space: stable by (src partition asc, gain desc); invalid moves last within a src
  std::vector<uint32_t> orig(n);
  for (uint32_t i = 0; i < n; ++i) orig[i] = i;
  auto valid = [&](uint32_t v) { return moves[v] != NONE; };
  std::stable_sort(orig.begin(), orig.end(), [&](uint32_t a, uint32_t b) {
    if (part[a] != part[b]) return part[a] < part[b];
    bool va = valid(a), vb = valid(b);
    if (va != vb) return va;
    if (!va) return false;
    return gains[a] > gains[b];
  });
  std::vector<uint32_t> src(n), dst(n), sz(n);
  std::vector<int64_t> w(n);
  std::vector<uint8_t> ok(n);
  for (uint32_t j = 0; j < n; ++j) {
    uint32_t v = orig[j];
    src[j] = part[v];
    dst[j] = moves[v];
    sz[j] = sizes[v];
    w[j] = gains[v];
    ok[j] = valid(v);
  }
  std::vector<uint32_t> ob(n), oe(n);
  for (uint32_t j = 0; j < n; ++j) {
    if (!ok[j]) {
      ob[j] = oe[j] = 0;
      continue;
    }
    ob[j] = (uint32_t)(std::lower_bound(src.begin(), src.end(), dst[j]) - src.begin());
    oe[j] = (uint32_t)(std::upper_bound(src.begin(), src.end(), dst[j]) - src.begin());
  }
  auto bucket = [](uint32_t s) -> int { return s == 0 ? 0 : 31 - __builtin_clz(s); };
  std::vector<uint32_t> next(n, NONE), prev(n, NONE), choice(n, NONE);
  const uint32_t WINDOW = 256;
  for (int it = 0; it < 4; ++it) {
    for (uint32_t i = 0; i < n; ++i) {
      choice[i] = NONE;
      if (next[i] != NONE || ob[i] >= oe[i]) continue;
      int bi = bucket(sz[i]);
      uint32_t lim = std::min(oe[i], ob[i] + WINDOW);
      uint32_t bj = NONE;
      int64_t bw = 0;
      uint32_t bd = 0;
      for (uint32_t j = ob[i]; j < lim; ++j) {
        if (j == i || prev[j] != NONE || !ok[j]) continue;
        if (std::abs(bi - bucket(sz[j])) > 1) continue;
        uint32_t d = sz[i] > sz[j] ? sz[i] - sz[j] : sz[j] - sz[i];
        if (bj == NONE || w[j] > bw || (w[j] == bw && d < bd) || (w[j] == bw && d == bd && j < bj)) {
          bj = j;
          bw = w[j];
          bd = d;
        }
      }
      choice[i] = bj;
    }
    // resolve: each successor keeps the claim with the best (score, -pred)
    std::vector<uint32_t> claim(n, NONE);
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t s = choice[i];
      if (s == NONE || prev[s] != NONE) continue;
      uint32_t c = claim[s];
      if (c == NONE) {
        claim[s] = i;
        continue;
      }
      uint32_t di = sz[i] > sz[s] ? sz[i] - sz[s] : sz[s] - sz[i];
      uint32_t dc = sz[c] > sz[s] ? sz[c] - sz[s] : sz[s] - sz[c];
      if (di < dc || (di == dc && i < c)) claim[s] = i;
    }
    for (uint32_t s = 0; s < n; ++s) {
      uint32_t p = claim[s];
      if (p == NONE || prev[s] != NONE) continue;
      prev[s] = p;
      if (next[p] == NONE) next[p] = s;
    }
  }
  // component (head, or min id on a cycle / long chain within 256 steps) and position
  std::vector<uint32_t> comp(n), pos(n);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t cur = i, mn = i;
    bool found = false;
    for (uint32_t d = 0; d < 256; ++d) {
      uint32_t p = prev[cur];
      if (p == NONE) {
        comp[i] = cur;
        pos[i] = d;
        found = true;
        break;
      }
      cur = p;
      mn = std::min(mn, cur);
    }
    if (!found) {
      comp[i] = mn;
      pos[i] = 0;
    }
  }
  // component weights; invalid moves form their own singleton components ranked last
  std::map<uint32_t, std::pair<int64_t, bool>> cw;  // comp -> (sum, all valid)
  for (uint32_t i = 0; i < n; ++i) {
    auto& c = cw[comp[i]];
    if (ok[i]) c.first += w[i];
    else c.second = true;  // contains an invalid move
  }
  std::vector<std::pair<uint32_t, std::pair<int64_t, bool>>> comps(cw.begin(), cw.end());
  std::stable_sort(comps.begin(), comps.end(), [](const auto& a, const auto& b) {
    if (a.second.second != b.second.second) return !a.second.second;
    if (!a.second.second && a.second.first != b.second.first) return a.second.first > b.second.first;
    return a.first < b.first;
  });
  std::unordered_map<uint32_t, uint32_t> crank;
  for (uint32_t r = 0; r < comps.size(); ++r) crank[comps[r].first] = r;
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    uint32_t ra = crank[comp[a]], rb = crank[comp[b]];
    if (ra != rb) return ra < rb;
    if (pos[a] != pos[b]) return pos[a] < pos[b];
    return a < b;
  });
  ranks.assign(n, 0);
  for (uint32_t r = 0; r < n; ++r) ranks[orig[order[r]]] = r;
  if (links) {
    links->prev.assign(n, NONE);
    links->comp.assign(n, NONE);
    links->pos.assign(n, 0);
    for (uint32_t j = 0; j < n; ++j) {
      links->prev[orig[j]] = prev[j] == NONE ? NONE : orig[prev[j]];
      links->comp[orig[j]] = orig[comp[j]];
      links->pos[orig[j]] = pos[j];
    }
  }
}

// gains (fm_refinement_gains_kernel, kernels/refinement.cu:82-209) in 2^18 fixed point:
// gain(q) = saving - sum_{e: c_e(q)=0} W_e; best q != p with psizes[q] + s - floor(s /
// discount) <= omega, by (gain, mix32(q + repeat)); moves[v] = NONE when no q fits.
void best_moves(const Level& L, bool incc, uint32_t omega, uint32_t discount, uint32_t k, uint32_t fm_repeat,
                const uint32_t* part, const uint32_t* psizes,
                const std::vector<std::unordered_map<uint32_t, uint32_t>>& cnt, std::vector<uint32_t>& moves,
                std::vector<int64_t>& gains) {
  const uint32_t n = L.n;
  moves.assign(n, NONE);
  gains.assign(n, 0);
  std::vector<int64_t> conn(k, 0);
  std::vector<uint32_t> touched;
  for (uint32_t v = 0; v < n; ++v) {
    uint32_t p = part[v], s = L.sizes[v];
    int64_t saving = 0, Wv = 0;
    touched.clear();
    for (uint64_t j = L.toff[v]; j < L.toff[v + 1]; ++j) {
      uint32_t e = L.tch[j];
      int64_t We = L.W[e];
      Wv += We;
      uint32_t c = cnt[e].count(p) ? cnt[e].at(p) : 0;
      if (incc) {
        if (c >= 1) saving += (int64_t)((uint64_t)L.W[e] / ((uint64_t)c * c));
      } else if (c == 1) {
        saving += We;
      }
      for (auto& kv : cnt[e]) {
        if (kv.first == p) continue;
        if (conn[kv.first] == 0) touched.push_back(kv.first);
        conn[kv.first] += We;
      }
    }
    uint32_t bq = NONE;
    int64_t bg = 0;
    uint32_t bh = 0;
    for (uint32_t q = 0; q < k; ++q) {
      if (q == p) continue;
      if ((uint32_t)(psizes[q] + s - s / discount) > omega) continue;
      int64_t g = saving - Wv + conn[q];
      uint32_t h = mix32(q + fm_repeat);
      if (bq == NONE || g > bg || (g == bg && h > bh)) {
        bq = q;
        bg = g;
        bh = h;
      }
    }
    for (uint32_t q : touched) conn[q] = 0;
    moves[v] = bq;
    gains[v] = bq == NONE ? 0 : bg;
  }
}

void analyze(const Level& L, int mode, uint32_t omega, uint32_t delta, uint32_t R, uint32_t k, uint32_t fm_repeat,
             bool chainup, const uint32_t* part, const uint32_t* psizes, Analysis& A) {
  const uint32_t n = L.n;
  const bool incc = mode == 0;
  const uint32_t discount = discount_of(fm_repeat, R);
  std::vector<std::unordered_map<uint32_t, uint32_t>> cnt, dcnt;
  hedge_counts(L, part, cnt, &dcnt);
  best_moves(L, incc, omega, discount, k, fm_repeat, part, psizes, cnt, A.moves, A.gains);
  // ranks
  if (chainup) {
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      bool va = A.moves[a] != NONE, vb = A.moves[b] != NONE;
      if (va != vb) return va;
      if (va && A.gains[a] != A.gains[b]) return A.gains[a] > A.gains[b];
      return a < b;
    });
    A.ranks.assign(n, 0);
    for (uint32_t r = 0; r < n; ++r) A.ranks[order[r]] = r;
  } else {
    chain_ranks(n, part, A.moves, A.gains, L.sizes, A.ranks);
  }
  A.V = 0;
  for (uint32_t v = 0; v < n; ++v) A.V += A.moves[v] != NONE;
  std::vector<uint32_t> by_rank(n);
  for (uint32_t v = 0; v < n; ++v) by_rank[A.ranks[v]] = v;
  // cascade (fm_refinement_cascade_kernel, kernels/refinement.cu:216-322): gain of each
  // move assuming every better-ranked move was applied; then prefix sums.
  A.seq.assign(n, 0);
  for (uint32_t r = 0; r < A.V; ++r) {
    uint32_t v = by_rank[r];
    uint32_t p = part[v], q = A.moves[v];
    int64_t sc = 0;
    for (uint64_t j = L.toff[v]; j < L.toff[v + 1]; ++j) {
      uint32_t e = L.tch[j];
      int64_t cp = cnt[e].count(p) ? cnt[e].at(p) : 0;
      int64_t cq = cnt[e].count(q) ? cnt[e].at(q) : 0;
      for (uint64_t i = L.hoff[e]; i < L.hoff[e + 1]; ++i) {
        uint32_t x = L.pins[i];
        if (A.moves[x] == NONE || A.ranks[x] >= r) continue;
        if (A.moves[x] == p) cp++;
        else if (A.moves[x] == q) cq++;
        if (part[x] == p) cp--;
        else if (part[x] == q) cq--;
      }
      if (incc) {
        if (cp >= 1) sc += (int64_t)((uint64_t)L.W[e] / ((uint64_t)cp * cp));
      } else if (cp == 1) {
        sc += L.W[e];
      }
      if (cq == 0) sc -= L.W[e];
    }
    A.seq[r] = sc;
  }
  for (uint32_t r = 1; r < A.V; ++r) A.seq[r] += A.seq[r - 1];
  // size validity (build/flag_size_events, kernels/refinement.cu:374-464)
  A.sviol.assign(n, 0);
  {
    std::vector<int64_t> cur(k);
    for (uint32_t q = 0; q < k; ++q) cur[q] = psizes[q];
    auto excess = [&](int64_t x) { return x > (int64_t)omega ? x - (int64_t)omega : 0; };
    for (uint32_t r = 0; r < A.V; ++r) {
      uint32_t v = by_rank[r];
      uint32_t p = part[v], q = A.moves[v];
      int64_t s = L.sizes[v];
      int64_t d = 0;
      d += excess(cur[p] - s) - excess(cur[p]);
      cur[p] -= s;
      d += excess(cur[q] + s) - excess(cur[q]);
      cur[q] += s;
      A.sviol[r] = (int32_t)d;
    }
    for (uint32_t r = 1; r < A.V; ++r) A.sviol[r] += A.sviol[r - 1];
  }
  // inbound validity (hedge events + inbound size events, :307-491); a no-op for k-way
  A.iviol.assign(n, 0);
  if (incc) {
    std::vector<int64_t> isz(k, 0);
    for (uint32_t e = 0; e < L.E; ++e)
      for (auto& kv : dcnt[e]) isz[kv.first]++;
    std::vector<std::unordered_map<uint32_t, int64_t>> dc(L.E);
    for (uint32_t e = 0; e < L.E; ++e)
      for (auto& kv : dcnt[e]) dc[e][kv.first] = kv.second;
    // A partition's first inbound-size change in the sequence is taken to start from a
    // valid state (flag_inbound_events_kernel, kernels/refinement.cu:643-644).
    auto invalid = [&](int64_t x) { return x > (int64_t)delta; };
    std::vector<uint8_t> touched(k, 0);
    for (uint32_t r = 0; r < A.V; ++r) {
      uint32_t v = by_rank[r];
      uint32_t p = part[v], q = A.moves[v];
      int32_t d = 0;
      for (uint64_t j = L.toff[v]; j < L.toff[v] + L.inb[v]; ++j) {
        uint32_t e = L.tch[j];
        int64_t& a = dc[e][p];
        int64_t before = a;
        a -= 1;
        if (before > 0 && a == 0) {
          bool was = touched[p] && invalid(isz[p]);
          touched[p] = 1;
          isz[p] -= 1;
          d += (int32_t)invalid(isz[p]) - (int32_t)was;
        }
        int64_t& b = dc[e][q];
        before = b;
        b += 1;
        if (before == 0 && b > 0) {
          bool was = touched[q] && invalid(isz[q]);
          touched[q] = 1;
          isz[q] += 1;
          d += (int32_t)invalid(isz[q]) - (int32_t)was;
        }
      }
      A.iviol[r] = d;
    }
    for (uint32_t r = 1; r < A.V; ++r) A.iviol[r] += A.iviol[r - 1];
  }
  // best_move_functor (headers/data_types.cuh:54-77): inbound-ok first, fewer size
  // violations, larger gain, earlier rank.
  A.best = NONE;
  for (uint32_t r = 0; r < A.V; ++r) {
    if (A.best == NONE) {
      A.best = r;
      continue;
    }
    uint32_t b = A.best;
    bool okr = A.iviol[r] == 0, okb = A.iviol[b] == 0;
    if (okr != okb) {
      if (okr) A.best = r;
      continue;
    }
    if (!okr) continue;
    if (A.sviol[r] != A.sviol[b]) {
      if (A.sviol[r] < A.sviol[b]) A.best = r;
      continue;
    }
    if (A.seq[r] > A.seq[b]) A.best = r;
  }
}

// One Jacobi label-propagation round, k-way (the reference's legacy "jacobi" step:
// fm_refinement_gains_kernel then jacobi_apply_moves_kernel, sources/init_part.cu:380-474,
// kernels/init_part.cu:89-151).  Every node proposes its best move with strict sizes
// (discount none) and tie hash mix32(q + round); proposals of positive gain are taken in
// (target asc, gain desc, node asc) order, and a move applies when the target's size before
// the round plus everything proposed into it up to and including this move fits omega
// (moves leaving the target are not counted on; once one fails, the rest of that target's
// list fails too).  The reference applies every proposal and breaks ties with coin flips;
// this rule is deterministic and keeps the sizes.  Returns the number of moves applied.
uint32_t jacobi_round(const Level& L, uint32_t omega, uint32_t k, uint32_t round, uint32_t* part, uint32_t* psizes) {
  std::vector<std::unordered_map<uint32_t, uint32_t>> cnt;
  hedge_counts(L, part, cnt, nullptr);
  std::vector<uint32_t> moves;
  std::vector<int64_t> gains;
  best_moves(L, false, omega, NONE, k, round, part, psizes, cnt, moves, gains);
  std::vector<uint32_t> movers;
  for (uint32_t v = 0; v < L.n; ++v)
    if (moves[v] != NONE && gains[v] > 0) movers.push_back(v);
  std::sort(movers.begin(), movers.end(), [&](uint32_t a, uint32_t b) {
    if (moves[a] != moves[b]) return moves[a] < moves[b];
    if (gains[a] != gains[b]) return gains[a] > gains[b];
    return a < b;
  });
  const std::vector<uint32_t> before(psizes, psizes + k);
  std::vector<uint64_t> entering(k, 0);
  uint32_t applied = 0;
  for (uint32_t v : movers) {
    const uint32_t q = moves[v], s = L.sizes[v];
    entering[q] += s;
    if ((uint64_t)before[q] + entering[q] > omega) continue;
    psizes[part[v]] -= s;
    psizes[q] += s;
    part[v] = q;
    ++applied;
  }
  return applied;
}

int64_t km1_fixed(const Level& L, const uint32_t* part) {
  int64_t total = 0;
  std::unordered_set<uint32_t> seen;
  for (uint32_t e = 0; e < L.E; ++e) {
    seen.clear();
    for (uint64_t i = L.hoff[e]; i < L.hoff[e + 1]; ++i) seen.insert(part[L.pins[i]]);
    if (!seen.empty()) total += (int64_t)L.W[e] * (int64_t)(seen.size() - 1);
  }
  return total;
}

// Label-propagation refinement (k-way, optional, before a level's FM rounds): up to
// `rounds` Jacobi rounds; a round that applies nothing ends it, and a round that raises
// lambda-1 is undone and ends it.
void lp_refine(const Level& L, uint32_t omega, uint32_t k, uint32_t rounds, uint32_t* part, uint32_t* psizes) {
  if (rounds == 0) return;
  int64_t cur = km1_fixed(L, part);
  std::vector<uint32_t> sp, ss;
  for (uint32_t r = 0; r < rounds; ++r) {
    sp.assign(part, part + L.n);
    ss.assign(psizes, psizes + k);
    if (jacobi_round(L, omega, k, r, part, psizes) == 0) break;
    const int64_t now = km1_fixed(L, part);
    if (now > cur) {
      std::copy(sp.begin(), sp.end(), part);
      std::copy(ss.begin(), ss.end(), psizes);
      break;
    }
    cur = now;
  }
}

// refinementRepeats loop with the reference's schedule (sources/refinement.cu:82-547).
void refine(const Level& L, int mode, uint32_t omega, uint32_t delta, uint32_t R, uint32_t k, bool final_counts,
            uint32_t* part, uint32_t* psizes, uint32_t* pinbound) {
  bool chainup = false;
  Analysis A;
  for (uint32_t rep = 0; rep < R; ++rep) {
    analyze(L, mode, omega, delta, R, k, rep, chainup, part, psizes, A);
    bool apply = false;
    int32_t sv = 1, iv = 1;
    int64_t g = 0;
    if (A.best != NONE) {
      sv = A.sviol[A.best];
      iv = A.iviol[A.best];
      g = A.seq[A.best];
      apply = sv <= 0 && iv <= 0 && g >= 0;
    }
    if (apply) {
      for (uint32_t v = 0; v < L.n; ++v) {
        if (A.moves[v] == NONE || A.ranks[v] > A.best) continue;
        psizes[part[v]] -= L.sizes[v];
        psizes[A.moves[v]] += L.sizes[v];
        part[v] = A.moves[v];
      }
    } else {
      if (sv > 0 && !chainup) chainup = true;
      else if (rep < R / 3) rep = R / 2;
      else if (rep < 2 * R / 3) rep = 2 * R / 3;
      else rep = R;
    }
  }
  if (final_counts && pinbound) {
    std::vector<std::unordered_map<uint32_t, uint32_t>> cnt, dcnt;
    hedge_counts(L, part, cnt, &dcnt);
    for (uint32_t q = 0; q < k; ++q) pinbound[q] = 0;
    for (uint32_t e = 0; e < L.E; ++e)
      for (auto& kv : dcnt[e]) pinbound[kv.first]++;
  }
}

// greedy consecutive packing shared by greedyMergeGroups and mergeSmallPartitions
// (sources/postprocess.cu:94-190): packs[pos] for items sorted by (size, inbound, id).
std::vector<uint32_t> greedy_packs(const std::vector<uint64_t>& s, const std::vector<uint64_t>& in, uint64_t omega,
                                   uint64_t delta) {
  size_t m = s.size();
  std::vector<uint64_t> ps(m + 1, 0), pi(m + 1, 0);
  for (size_t i = 0; i < m; ++i) {
    ps[i + 1] = ps[i] + s[i];
    pi[i + 1] = pi[i] + in[i];
  }
  std::vector<uint32_t> pack(m);
  size_t pos = 0;
  uint32_t id = 0;
  while (pos < m) {
    size_t us = std::upper_bound(ps.begin(), ps.end(), ps[pos] + omega) - ps.begin();
    size_t ui = std::upper_bound(pi.begin(), pi.end(), pi[pos] + delta) - pi.begin();
    size_t nxt = std::max(pos + 1, std::min(us, ui) - 1);
    for (size_t j = pos; j < nxt; ++j) pack[j] = id;
    id++;
    pos = nxt;
  }
  return pack;
}

void greedy_merge_groups(uint32_t n, uint32_t ng, const uint32_t* sizes, const uint32_t* inb, uint32_t omega,
                         uint32_t delta, uint32_t* groups, uint32_t* gsizes, uint32_t* new_ng) {
  if (n == 0 || ng < 2) {
    *new_ng = ng;
    return;
  }
  std::vector<uint64_t> gin(ng, 0);
  for (uint32_t v = 0; v < n; ++v) gin[groups[v]] += inb[v];
  std::vector<uint32_t> idx(ng);
  for (uint32_t g = 0; g < ng; ++g) idx[g] = g;
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
    if (gsizes[a] != gsizes[b]) return gsizes[a] < gsizes[b];
    if ((uint32_t)gin[a] != (uint32_t)gin[b]) return (uint32_t)gin[a] < (uint32_t)gin[b];
    return a < b;
  });
  std::vector<uint64_t> s(ng), in(ng);
  for (uint32_t i = 0; i < ng; ++i) {
    s[i] = gsizes[idx[i]];
    in[i] = (uint32_t)gin[idx[i]];
  }
  auto pack = greedy_packs(s, in, omega, delta);
  std::vector<uint32_t> rep(ng, NONE);
  for (uint32_t i = 0; i < ng; ++i) rep[pack[i]] = std::min(rep[pack[i]], idx[i]);
  std::vector<uint32_t> gmap(ng);
  for (uint32_t i = 0; i < ng; ++i) gmap[idx[i]] = rep[pack[i]];
  for (uint32_t v = 0; v < n; ++v) groups[v] = gmap[groups[v]];
  // zero-base in ascending representative id order
  std::vector<uint32_t> flag(ng, 0);
  for (uint32_t v = 0; v < n; ++v) flag[groups[v]] = 1;
  std::vector<uint32_t> nid(ng, 0);
  uint32_t c = 0;
  for (uint32_t g = 0; g < ng; ++g) {
    nid[g] = c;
    c += flag[g];
  }
  for (uint32_t g = 0; g < c; ++g) gsizes[g] = 0;
  for (uint32_t v = 0; v < n; ++v) {
    groups[v] = nid[groups[v]];
    gsizes[groups[v]] += sizes[v];
  }
  *new_ng = c;
}

void merge_small(uint32_t n, uint32_t k, const uint32_t* psizes, const uint32_t* pinb, uint32_t omega, uint32_t delta,
                 uint32_t* part) {
  std::vector<uint32_t> small;
  for (uint32_t p = 0; p < k; ++p)
    if (psizes[p] < 16) small.push_back(p);
  if (small.empty()) return;
  std::stable_sort(small.begin(), small.end(), [&](uint32_t a, uint32_t b) {
    if (psizes[a] != psizes[b]) return psizes[a] < psizes[b];
    if (pinb[a] != pinb[b]) return pinb[a] < pinb[b];
    return a < b;
  });
  std::vector<uint64_t> s(small.size()), in(small.size());
  for (size_t i = 0; i < small.size(); ++i) {
    s[i] = psizes[small[i]];
    in[i] = pinb[small[i]];
  }
  auto pack = greedy_packs(s, in, omega, delta);
  std::vector<uint32_t> rep(small.size(), NONE);
  for (size_t i = 0; i < small.size(); ++i) rep[pack[i]] = std::min(rep[pack[i]], small[i]);
  std::vector<uint32_t> pmap(k);
  for (uint32_t p = 0; p < k; ++p) pmap[p] = p;
  for (size_t i = 0; i < small.size(); ++i) pmap[small[i]] = rep[pack[i]];
  for (uint32_t v = 0; v < n; ++v) part[v] = pmap[part[v]];
}

uint32_t relabel(uint32_t n, uint32_t* part) {
  std::vector<uint32_t> ids(part, part + n);
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  std::unordered_map<uint32_t, uint32_t> m;
  for (uint32_t i = 0; i < ids.size(); ++i) m[ids[i]] = i;
  for (uint32_t v = 0; v < n; ++v) part[v] = m[part[v]];
  return (uint32_t)ids.size();
}

void coarsen_hedges(uint32_t E, const uint64_t* off, const uint32_t* pins, const uint32_t* srcc, const uint32_t* groups,
                    std::vector<uint64_t>& coff, std::vector<uint32_t>& cpins, std::vector<uint32_t>& csrcc) {
  coff.assign(E + 1, 0);
  cpins.clear();
  csrcc.assign(E, 0);
  std::vector<uint32_t> d, s;
  for (uint32_t e = 0; e < E; ++e) {
    d.clear();
    s.clear();
    for (uint64_t i = off[e]; i < off[e + 1]; ++i) {
      uint32_t g = groups[pins[i]];
      if (i - off[e] < srcc[e]) s.push_back(g);
      else d.push_back(g);
    }
    std::sort(d.begin(), d.end());
    d.erase(std::unique(d.begin(), d.end()), d.end());
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    std::vector<uint32_t> s2;
    for (uint32_t g : s)
      if (!std::binary_search(d.begin(), d.end(), g)) s2.push_back(g);
    csrcc[e] = (uint32_t)s2.size();
    cpins.insert(cpins.end(), s2.begin(), s2.end());
    cpins.insert(cpins.end(), d.rbegin(), d.rend());
    coff[e + 1] = cpins.size();
  }
}

void build_level(Level& L, uint32_t n, uint32_t E, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                 const uint64_t* toff, const uint32_t* tch, const uint32_t* inb, const float* w,
                 const uint32_t* sizes) {
  L.n = n;
  L.E = E;
  L.hoff = hoff;
  L.pins = pins;
  L.srcc = srcc;
  L.toff = toff;
  L.tch = tch;
  L.inb = inb;
  L.sizes = sizes;
  L.W.resize(E);
  for (uint32_t e = 0; e < E; ++e) L.W[e] = fixed_weight(w[e]);
}

// The initial partitioner's input (sources/init_part.cu:749-888).  Pruning
// (armonic_degree_score_kernel and prune_hedges_kernel, kernels/init_part.cu:221-281): a
// hedge's score is its weight over the sum of its pins' inverse degrees, about f * N * floor(P/E)
// hedges are kept (f = 1 is the reference's target), kept weights are divided by the keep probability.  The reference draws
// with cuRAND Philox in float; the rule restated here is the build's exact one:
//   inv_e = sum_{v in e} floor(2^24 / deg(v)); ratio_e = min(floor(W_e 2^24 / inv_e), floor((2^64-1)/E))
//   p_e = S ? min(1, T * ratio_e / S) : 1 in IEEE double, T = f N floor(P/E), S = sum ratio_e
//   keep iff (splitmix64(86 ^ e * 0x9E3779B97F4A7C15) >> 11) < p_e * 2^53; W'_e = min(floor(W_e / p_e), 2^40)
// Then the file writer's filter (:822-847): hedges of >= 2 pins, pins sorted, identical
// pin sets merged with their weights summed (kept here in order of first appearance).
struct IpInput {
  std::vector<uint64_t> off{0};
  std::vector<uint32_t> pins;
  std::vector<uint64_t> w;
};

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

std::vector<uint64_t> prune_weights(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const float* w,
                                    uint32_t mult) {
  std::vector<uint64_t> out(e, 0);
  if (e == 0) return out;
  const uint64_t P = hoff[e] - hoff[0];
  std::vector<uint32_t> deg(n, 0);
  for (uint64_t i = hoff[0]; i < hoff[e]; ++i) deg[pins[i]]++;
  std::vector<uint64_t> ratio(e, 0);
  const uint64_t cap = ~0ull / e;
  uint64_t S = 0;
  for (uint32_t x = 0; x < e; ++x) {
    uint64_t inv = 0;
    for (uint64_t i = hoff[x]; i < hoff[x + 1]; ++i) inv += 16777216u / deg[pins[i]];
    if (inv) ratio[x] = std::min<uint64_t>(((uint64_t)fixed_weight(w[x]) << 24) / inv, cap);
    S += ratio[x];
  }
  const double T = (double)((uint64_t)mult * n * (P / e)), Sd = (double)S;
  for (uint32_t x = 0; x < e; ++x) {
    volatile double prod = T * (double)ratio[x];  // rounded on its own, as on the device
    const double p = S == 0 ? 1.0 : std::min(1.0, prod / Sd);
    const uint64_t u = splitmix64(86ull ^ ((uint64_t)x * 0x9E3779B97F4A7C15ull)) >> 11;
    if ((double)u < p * 9007199254740992.0) {
      const double sc = (double)fixed_weight(w[x]) / p;
      out[x] = sc >= 1099511627776.0 ? 1099511627776ull : (uint64_t)sc;
    }
  }
  return out;
}

int g_ip_threads = 0;  // 1: the initial partitioner runs on one thread; 0: its default

IpInput ip_input(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const float* w, uint32_t prune) {
  IpInput in;
  std::vector<uint64_t> pw;
  if (prune) pw = prune_weights(n, e, hoff, pins, w, prune);
  std::map<std::vector<uint32_t>, uint32_t> first;
  for (uint32_t x = 0; x < e; ++x) {
    if (prune && pw[x] == 0) continue;
    std::vector<uint32_t> set(pins + hoff[x], pins + hoff[x + 1]);
    std::sort(set.begin(), set.end());
    set.erase(std::unique(set.begin(), set.end()), set.end());
    if (set.size() < 2) continue;
    const uint64_t wx = prune ? pw[x] : fixed_weight(w[x]);
    auto it = first.find(set);
    if (it != first.end()) {
      in.w[it->second] += wx;
      continue;
    }
    first.emplace(set, (uint32_t)in.w.size());
    in.pins.insert(in.pins.end(), set.begin(), set.end());
    in.off.push_back(in.pins.size());
    in.w.push_back(wx);
  }
  return in;
}

// Best-fit-decreasing start of the device initial partitioner (modelled on the reference's
// legacy host BFD packing, sources/init_part.cu:37-180): the nodes in a seeded shuffle
// (Fisher-Yates over a splitmix64 stream started at seed * 0x9E3779B97F4A7C15 + 1), stably
// sorted by size descending; each goes to the fullest bin it fits in below the average load
// A = min(omega, ceil(total / k)), else to the fullest bin it fits in below omega, else to
// the lightest bin (ties: lower bin in all three).
void bfd_start(const std::vector<uint32_t>& sizes, uint32_t k, uint32_t omega, uint64_t seed, std::vector<uint32_t>& part,
               std::vector<uint32_t>& loads) {
  const uint32_t n = (uint32_t)sizes.size();
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
  for (uint32_t i = n; i >= 2; --i) {
    st = splitmix64(st);
    std::swap(order[i - 1], order[st % i]);
  }
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return sizes[a] > sizes[b]; });
  part.assign(n, 0);
  loads.assign(k, 0);
  uint64_t total = 0;
  for (uint32_t s : sizes) total += s;
  const uint64_t avg = std::min<uint64_t>(omega, (total + k - 1) / k);
  for (uint32_t v : order) {
    uint32_t pick = NONE;
    for (uint64_t cap : {avg, (uint64_t)omega}) {
      for (uint32_t b = 0; b < k; ++b) {
        if ((uint64_t)loads[b] + sizes[v] > cap) continue;
        if (pick == NONE || loads[b] > loads[pick]) pick = b;
      }
      if (pick != NONE) break;
    }
    if (pick == NONE) {
      pick = 0;
      for (uint32_t b = 1; b < k; ++b)
        if (loads[b] < loads[pick]) pick = b;
    }
    part[v] = pick;
    loads[pick] += sizes[v];
  }
}

// The device initial partitioner (k-way), restated: the level's hedges merged (identical
// pin sets summed, fewer than two pins dropped, optional pruning), then coarsened further by
// the pipeline's own stages (candidates with cnc 16 and no inbound limit, pairing, coarse
// hedges, merged again) while it has more than max(16 k, 256) nodes, fewer than 32 levels
// and the pairing shrinks it below 95%; on the coarsest merged level, `starts` BFD starts each
// improved by up to `rounds` Jacobi rounds (until one applies nothing), the start of lowest
// lambda-1 kept (the first on ties); then projected back level by level with `repeats` k-way
// refinement rounds each.  The coarsening target is max(per_part * k, 256) nodes.  Merged weights are the summed 2^18 fixed-point weights over 2^18
// as float.
struct MiniLevel {
  uint32_t n = 0, E = 0;
  std::vector<uint64_t> hoff, toff;
  std::vector<uint32_t> pins, srcc, tch, inb, sizes, groups;
  std::vector<float> w;
};

void mini_level(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const float* w, uint32_t prune,
                std::vector<uint32_t> sizes, MiniLevel& M) {
  IpInput in = ip_input(n, e, hoff, pins, w, prune);
  M.n = n;
  M.E = (uint32_t)in.w.size();
  M.hoff = std::move(in.off);
  M.pins = std::move(in.pins);
  M.srcc.assign(M.E, 0);
  M.w.resize(M.E);
  for (uint32_t x = 0; x < M.E; ++x) M.w[x] = (float)((double)in.w[x] / 262144.0);
  M.sizes = std::move(sizes);
  touching(n, M.E, M.hoff.data(), M.pins.data(), M.srcc.data(), M.toff, M.tch, M.inb);
}

void device_ip(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const float* w,
               const uint32_t* sizes, uint32_t k, uint32_t omega, uint32_t prune, uint32_t starts, uint32_t rounds,
               uint32_t per_part, uint32_t repeats, uint32_t* part_out, uint32_t* psizes_out) {
  k = std::max<uint32_t>(k, 1);
  const uint32_t target = std::max<uint32_t>(std::max<uint32_t>(per_part, 1) * k, 256), cnc = 16,
                 delta = 0x7FFFFFFFu;
  std::vector<MiniLevel> lv(1);
  mini_level(n, e, hoff, pins, w, prune, std::vector<uint32_t>(sizes, sizes + n), lv[0]);
  while (lv.back().n > target && lv.size() < 32) {
    MiniLevel& M = lv.back();
    std::vector<uint32_t> pairs((uint64_t)M.n * cnc), scores((uint64_t)M.n * cnc);
    orc_candidates(M.n, M.E, M.hoff.data(), M.pins.data(), M.srcc.data(), M.toff.data(), M.tch.data(), M.inb.data(),
                   M.w.data(), M.sizes.data(), cnc, omega, delta, pairs.data(), scores.data());
    uint32_t nn = 0;
    std::vector<uint32_t> groups(M.n), gs(M.n), ung(M.n);
    std::vector<uint64_t> uoff(M.n + 1);
    orc_group_nodes(M.n, M.inb.data(), pairs.data(), scores.data(), M.sizes.data(), cnc, omega, delta, &nn,
                    groups.data(), gs.data(), ung.data(), uoff.data(), nullptr);
    if ((float)nn / (float)M.n > 0.95f) break;
    gs.resize(nn);
    std::vector<uint64_t> co;
    std::vector<uint32_t> cp, cs;
    coarsen_hedges(M.E, M.hoff.data(), M.pins.data(), M.srcc.data(), groups.data(), co, cp, cs);
    M.groups = std::move(groups);
    MiniLevel N;
    mini_level(nn, M.E, co.data(), cp.data(), M.w.data(), 0, std::move(gs), N);
    lv.push_back(std::move(N));
  }
  std::vector<uint32_t> part, psizes;
  {
    const MiniLevel& M = lv.back();
    Level L;
    build_level(L, M.n, M.E, M.hoff.data(), M.pins.data(), M.srcc.data(), M.toff.data(), M.tch.data(), M.inb.data(),
                M.w.data(), M.sizes.data());
    int64_t best = INT64_MAX;
    std::vector<uint32_t> sp, ss;
    for (uint32_t s = 0; s < std::max<uint32_t>(starts, 1); ++s) {
      bfd_start(M.sizes, k, omega, s, sp, ss);
      for (uint32_t r = 0; r < rounds; ++r)
        if (jacobi_round(L, omega, k, r, sp.data(), ss.data()) == 0) break;
      const int64_t km1 = km1_fixed(L, sp.data());
      if (km1 < best) {
        best = km1;
        part = sp;
        psizes = ss;
      }
    }
  }
  for (size_t d = lv.size() - 1; d-- > 0;) {
    const MiniLevel& M = lv[d];
    std::vector<uint32_t> fine(M.n);
    for (uint32_t v = 0; v < M.n; ++v) fine[v] = part[M.groups[v]];
    part.swap(fine);
    Level L;
    build_level(L, M.n, M.E, M.hoff.data(), M.pins.data(), M.srcc.data(), M.toff.data(), M.tch.data(), M.inb.data(),
                M.w.data(), M.sizes.data());
    refine(L, 1, omega, delta, repeats, k, false, part.data(), psizes.data(), nullptr);
  }
  std::copy(part.begin(), part.end(), part_out);
  std::copy(psizes.begin(), psizes.end(), psizes_out);
}

}  // namespace

// =======================================================================================
extern "C" {

void* orc_hg_load(const char* path, int dedup) {
  HG* hg = new HG();
  std::string p(path);
  bool ok = false;
  if (ends_with(p, ".hgr")) ok = load_hgr(p, *hg);
  else if (ends_with(p, ".snn")) ok = load_bin(p, false, *hg);
  else if (ends_with(p, ".axh")) ok = load_bin(p, true, *hg);
  if (!ok) {
    delete hg;
    return nullptr;
  }
  if (dedup) {
    uint64_t np = dedup_rule2(hg->e(), hg->off.data(), hg->pins.data(), hg->srcc.data());
    hg->pins.resize(np);
  }
  return hg;
}

void orc_hg_sizes(const void* h, uint32_t* n, uint32_t* e, uint64_t* p) {
  const HG* hg = (const HG*)h;
  *n = hg->n;
  *e = hg->e();
  *p = hg->pins.size();
}

void orc_hg_copy(const void* h, uint64_t* hoff, uint32_t* pins, uint32_t* srcc, float* w) {
  const HG* hg = (const HG*)h;
  std::copy(hg->off.begin(), hg->off.end(), hoff);
  std::copy(hg->pins.begin(), hg->pins.end(), pins);
  std::copy(hg->srcc.begin(), hg->srcc.end(), srcc);
  std::copy(hg->w.begin(), hg->w.end(), w);
}

void orc_hg_free(void* h) { delete (HG*)h; }

uint64_t orc_dedup(uint32_t e, uint64_t* hoff, uint32_t* pins, uint32_t* srcc) {
  return dedup_rule2(e, hoff, pins, srcc);
}

uint32_t orc_fixed_weight(float w) { return fixed_weight(w); }
uint32_t orc_noise(uint32_t a, uint32_t b) { return noise64(a, b); }
uint32_t orc_hash_uint32(uint32_t x) { return mix32(x); }

int orc_build_touching(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                       uint64_t* toff, uint32_t* tch, uint32_t* inb) {
  std::vector<uint64_t> o;
  std::vector<uint32_t> t, ib;
  touching(n, e, hoff, pins, srcc, o, t, ib);
  std::copy(o.begin(), o.end(), toff);
  std::copy(t.begin(), t.end(), tch);
  std::copy(ib.begin(), ib.end(), inb);
  return 0;
}

// buildNeighbors (sources/construction.cu:242-340): N(v) = union of touching hedges'
// pins minus v (kernels/construction.cu:157-243), ascending.
int orc_build_neighbors(uint32_t n, const uint64_t* hoff, const uint32_t* pins, const uint64_t* toff,
                        const uint32_t* tch, uint64_t* noff, uint32_t* nbr) {
  std::vector<uint32_t> tmp;
  noff[0] = 0;
  for (uint32_t v = 0; v < n; ++v) {
    tmp.clear();
    for (uint64_t j = toff[v]; j < toff[v + 1]; ++j) {
      uint32_t e = tch[j];
      for (uint64_t i = hoff[e]; i < hoff[e + 1]; ++i)
        if (pins[i] != v) tmp.push_back(pins[i]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    if (nbr) std::copy(tmp.begin(), tmp.end(), nbr + noff[v]);
    noff[v + 1] = noff[v] + tmp.size();
  }
  return 0;
}

// candidatesProposal (kernels/coarsening.cu:10-213): score(u,v) = noise(u,v) +
// sum_{e containing u and v} floor(W_e / |e|) (u32 wrap); v is a candidate when
// size(u)+size(v) <= omega and |in(u) U in(v)| <= delta and its score is non-zero; the
// best `cnc` by (score desc, id desc) are kept, empty slots UINT32_MAX / 0.
int orc_candidates(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                   const uint64_t* toff, const uint32_t* tch, const uint32_t* inb, const float* w,
                   const uint32_t* sizes, uint32_t cnc, uint32_t omega, uint32_t delta, uint32_t* pairs,
                   uint32_t* scores) {
  (void)e;
  // dense accumulators over v plus the list of touched v (order-independent integer sums)
  std::vector<uint32_t> acc_score(n, 0), acc_inter(n, 0);
  std::vector<uint8_t> seen(n, 0);
  std::vector<uint32_t> touched;
  std::vector<std::pair<uint32_t, uint32_t>> cand;
  for (uint32_t u = 0; u < n; ++u) {
    touched.clear();
    for (uint64_t j = toff[u]; j < toff[u + 1]; ++j) {
      uint32_t h = tch[j];
      bool inbound = (j - toff[u]) < inb[u];
      uint64_t sz = hoff[h + 1] - hoff[h];
      uint32_t contrib = (uint32_t)((uint64_t)fixed_weight(w[h]) / sz);
      for (uint64_t i = hoff[h]; i < hoff[h + 1]; ++i) {
        uint32_t v = pins[i];
        if (v == u) continue;
        if (!seen[v]) {
          seen[v] = 1;
          touched.push_back(v);
        }
        acc_score[v] += contrib;
        if (inbound && (i - hoff[h]) >= srcc[h]) acc_inter[v] += 1;
      }
    }
    cand.clear();
    for (uint32_t v : touched) {
      const uint32_t sc = acc_score[v], in = acc_inter[v];
      seen[v] = 0;
      acc_score[v] = 0;
      acc_inter[v] = 0;
      if ((uint32_t)(sizes[u] + sizes[v]) > omega) continue;
      if ((uint32_t)((inb[v] - in) + inb[u]) > delta) continue;
      uint32_t s = sc + noise64(u, v);
      if (s == 0) continue;
      cand.push_back({s, v});
    }
    // top-cnc by (score desc, id desc); ids are distinct, so the order is total
    const size_t top = std::min<size_t>(cnc, cand.size());
    std::partial_sort(cand.begin(), cand.begin() + top, cand.end(), [](const auto& a, const auto& b) {
      return a.first != b.first ? a.first > b.first : a.second > b.second;
    });
    for (uint32_t i = 0; i < cnc; ++i) {
      pairs[(uint64_t)u * cnc + i] = i < cand.size() ? cand[i].second : NONE;
      scores[(uint64_t)u * cnc + i] = i < cand.size() ? cand[i].first : 0u;
    }
  }
  return 0;
}

int orc_group_nodes(uint32_t n, const uint32_t* inb, const uint32_t* pairs, const uint32_t* scores,
                    const uint32_t* sizes, uint32_t cnc, uint32_t omega, uint32_t delta, uint32_t* new_n,
                    uint32_t* groups, uint32_t* group_sizes, uint32_t* ungroups, uint64_t* uoff,
                    uint32_t* raw_groups) {
  std::vector<uint32_t> raw;
  pairing(n, pairs, scores, cnc, raw);
  orphans(n, pairs, cnc, sizes, inb, omega, delta, raw);
  if (raw_groups) std::copy(raw.begin(), raw.end(), raw_groups);
  // relabel: zero-based in ascending raw-id order (sources/coarsening.cu:152-207)
  std::vector<uint32_t> flag(n, 0), nid(n, 0);
  for (uint32_t v = 0; v < n; ++v) flag[raw[v]] = 1;
  uint32_t c = 0;
  for (uint32_t g = 0; g < n; ++g) {
    nid[g] = c;
    c += flag[g];
  }
  *new_n = c;
  for (uint32_t g = 0; g < c; ++g) group_sizes[g] = 0;
  std::vector<uint32_t> cntg(c + 1, 0);
  for (uint32_t v = 0; v < n; ++v) {
    groups[v] = nid[raw[v]];
    group_sizes[groups[v]] += sizes[v];
    cntg[groups[v] + 1]++;
  }
  uoff[0] = 0;
  for (uint32_t g = 0; g < c; ++g) uoff[g + 1] = uoff[g] + cntg[g + 1];
  std::vector<uint64_t> pos(uoff, uoff + c);
  for (uint32_t v = 0; v < n; ++v) ungroups[pos[groups[v]]++] = v;
  return 0;
}

int orc_coarsen_hedges(uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                       const uint32_t* groups, uint64_t* choff, uint32_t* cpins, uint32_t* csrcc) {
  std::vector<uint64_t> o;
  std::vector<uint32_t> p, s;
  coarsen_hedges(e, hoff, pins, srcc, groups, o, p, s);
  std::copy(o.begin(), o.end(), choff);
  if (cpins) std::copy(p.begin(), p.end(), cpins);
  std::copy(s.begin(), s.end(), csrcc);
  return 0;
}

// coarsenNeighbors (sources/construction.cu:342-561; kernels/construction.cu:495-605).
int orc_coarsen_neighbors(uint32_t n, const uint64_t* noff, const uint32_t* nbr, const uint32_t* groups,
                          uint32_t new_n, uint64_t* cnoff, uint32_t* cnbr) {
  std::vector<std::vector<uint32_t>> sets(new_n);
  for (uint32_t v = 0; v < n; ++v) {
    uint32_t g = groups[v];
    for (uint64_t i = noff[v]; i < noff[v + 1]; ++i) {
      uint32_t h = groups[nbr[i]];
      if (h != g) sets[g].push_back(h);
    }
  }
  cnoff[0] = 0;
  for (uint32_t g = 0; g < new_n; ++g) {
    auto& s = sets[g];
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    if (cnbr) std::copy(s.begin(), s.end(), cnbr + cnoff[g]);
    cnoff[g + 1] = cnoff[g] + s.size();
  }
  return 0;
}

int orc_refine_analyze(int mode, uint32_t omega, uint32_t delta, uint32_t refine_repeats, uint32_t n, uint32_t e,
                       const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc, const uint64_t* toff,
                       const uint32_t* tch, const uint32_t* inb, const float* w, const uint32_t* sizes, uint32_t k,
                       uint32_t fm_repeat, int chainup, const uint32_t* part, const uint32_t* psizes,
                       uint32_t* moves, int64_t* gains, uint32_t* ranks, int64_t* seq_gains, int32_t* size_viol,
                       int32_t* inbound_viol, uint32_t* best_rank) {
  Level L;
  build_level(L, n, e, hoff, pins, srcc, toff, tch, inb, w, sizes);
  Analysis A;
  analyze(L, mode, omega, delta, refine_repeats, k, fm_repeat, chainup != 0, part, psizes, A);
  std::copy(A.moves.begin(), A.moves.end(), moves);
  std::copy(A.gains.begin(), A.gains.end(), gains);
  std::copy(A.ranks.begin(), A.ranks.end(), ranks);
  std::copy(A.seq.begin(), A.seq.end(), seq_gains);
  std::copy(A.sviol.begin(), A.sviol.end(), size_viol);
  std::copy(A.iviol.begin(), A.iviol.end(), inbound_viol);
  *best_rank = A.best;
  return 0;
}

int orc_refine(int mode, uint32_t omega, uint32_t delta, uint32_t refine_repeats, uint32_t n, uint32_t e,
               const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc, const uint64_t* toff,
               const uint32_t* tch, const uint32_t* inb, const float* w, const uint32_t* sizes, uint32_t k,
               int final_counts, uint32_t* part, uint32_t* psizes, uint32_t* pinbound) {
  Level L;
  build_level(L, n, e, hoff, pins, srcc, toff, tch, inb, w, sizes);
  refine(L, mode, omega, delta, refine_repeats, k, final_counts != 0, part, psizes, pinbound);
  return 0;
}

// inbound_sets_size_kernel (kernels/refinement.cu:1176-1219): hedges with at least one
// destination pin in the partition.
int orc_inbound_sizes(uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                      const uint32_t* part, uint32_t k, uint32_t* out) {
  for (uint32_t q = 0; q < k; ++q) out[q] = 0;
  std::unordered_set<uint32_t> seen;
  for (uint32_t h = 0; h < e; ++h) {
    seen.clear();
    for (uint64_t i = hoff[h] + srcc[h]; i < hoff[h + 1]; ++i)
      if (seen.insert(part[pins[i]]).second) out[part[pins[i]]]++;
  }
  return 0;
}

int64_t orc_connectivity_fixed(uint32_t e, const uint64_t* hoff, const uint32_t* pins, const float* w,
                               const uint32_t* part) {
  int64_t total = 0;
  std::unordered_set<uint32_t> seen;
  for (uint32_t h = 0; h < e; ++h) {
    seen.clear();
    for (uint64_t i = hoff[h]; i < hoff[h + 1]; ++i) seen.insert(part[pins[i]]);
    if (!seen.empty()) total += (int64_t)fixed_weight(w[h]) * (int64_t)(seen.size() - 1);
  }
  return total;
}

int orc_greedy_merge_groups(uint32_t n, uint32_t ngroups, const uint32_t* sizes, const uint32_t* inb, uint32_t omega,
                            uint32_t delta, uint32_t* groups, uint32_t* group_sizes, uint32_t* new_ngroups) {
  greedy_merge_groups(n, ngroups, sizes, inb, omega, delta, groups, group_sizes, new_ngroups);
  return 0;
}

int orc_merge_small_partitions(uint32_t n, uint32_t k, const uint32_t* psizes, const uint32_t* pinbound,
                               uint32_t omega, uint32_t delta, uint32_t* part) {
  merge_small(n, k, psizes, pinbound, omega, delta, part);
  return 0;
}

// The chaining step alone on given moves and gains (invalid moves: moves[v] == NONE):
// ranks[v] and the links behind them (prev[v], component head comp[v], position pos[v]).
int orc_chain_links(uint32_t n, const uint32_t* part, const uint32_t* moves, const int64_t* gains,
                    const uint32_t* sizes, uint32_t* ranks, uint32_t* prev, uint32_t* comp, uint32_t* pos) {
  std::vector<uint32_t> mv(moves, moves + n), rk;
  std::vector<int64_t> gn(gains, gains + n);
  ChainLinks L;
  chain_ranks(n, part, mv, gn, sizes, rk, &L);
  std::copy(rk.begin(), rk.end(), ranks);
  std::copy(L.prev.begin(), L.prev.end(), prev);
  std::copy(L.comp.begin(), L.comp.end(), comp);
  std::copy(L.pos.begin(), L.pos.end(), pos);
  return 0;
}

int orc_ip_input(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const float* w,
                 const uint32_t* sizes, int prune, uint32_t* e_out, uint64_t* p_out, uint64_t* off_out,
                 uint32_t* pins_out, uint64_t* w_out, uint64_t* vw_out) {
  IpInput in = ip_input(n, e, hoff, pins, w, (uint32_t)std::max(prune, 0));
  *e_out = (uint32_t)(in.off.size() - 1);
  *p_out = in.pins.size();
  if (off_out) std::copy(in.off.begin(), in.off.end(), off_out);
  if (pins_out) std::copy(in.pins.begin(), in.pins.end(), pins_out);
  if (w_out) std::copy(in.w.begin(), in.w.end(), w_out);
  if (vw_out)
    for (uint32_t v = 0; v < n; ++v) vw_out[v] = sizes[v];
  return 0;
}

void orc_set_ip_threads(int threads) { g_ip_threads = threads; }

int orc_jacobi_round(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint64_t* toff,
                     const uint32_t* tch, const uint32_t* inb, const float* w, const uint32_t* sizes, uint32_t k,
                     uint32_t omega, uint32_t round, uint32_t* part, uint32_t* psizes, uint32_t* applied) {
  std::vector<uint32_t> srcc(e, 0);
  Level L;
  build_level(L, n, e, hoff, pins, srcc.data(), toff, tch, inb, w, sizes);
  *applied = jacobi_round(L, omega, k, round, part, psizes);
  return 0;
}

int orc_lp_refine(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint64_t* toff,
                  const uint32_t* tch, const uint32_t* inb, const float* w, const uint32_t* sizes, uint32_t k,
                  uint32_t omega, uint32_t rounds, uint32_t* part, uint32_t* psizes) {
  std::vector<uint32_t> srcc(e, 0);
  Level L;
  build_level(L, n, e, hoff, pins, srcc.data(), toff, tch, inb, w, sizes);
  lp_refine(L, omega, k, rounds, part, psizes);
  return 0;
}

int orc_bfd_start(uint32_t n, const uint32_t* sizes, uint32_t k, uint32_t omega, uint64_t seed, uint32_t* part,
                  uint32_t* loads) {
  std::vector<uint32_t> s(sizes, sizes + n), p, l;
  bfd_start(s, k, omega, seed, p, l);
  std::copy(p.begin(), p.end(), part);
  std::copy(l.begin(), l.end(), loads);
  return 0;
}

int orc_device_initial_partition(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const float* w,
                                 const uint32_t* sizes, uint32_t k, uint32_t omega, int prune, uint32_t starts,
                                 uint32_t rounds, uint32_t per_part, uint32_t repeats, uint32_t* part,
                                 uint32_t* psizes) {
  device_ip(n, e, hoff, pins, w, sizes, k, omega, (uint32_t)std::max(prune, 0), starts, rounds, per_part, repeats, part,
            psizes);
  return 0;
}

int orc_relabel(uint32_t n, uint32_t* part, uint32_t* new_k) {
  *new_k = relabel(n, part);
  return 0;
}

int orc_initial_partition(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const float* w,
                          const uint32_t* sizes, uint32_t k, float eps, uint32_t omega, int prune, uint32_t* part,
                          uint32_t* psizes) {
  IpInput in = ip_input(n, e, hoff, pins, w, (uint32_t)std::max(prune, 0));
  hgp_ip::Hypergraph h;
  h.n = n;
  h.off = std::move(in.off);
  h.pins = std::move(in.pins);
  h.w = std::move(in.w);
  h.vw.assign(sizes, sizes + n);
  // incidence: the hedges of every node in ascending order
  h.ioff.assign((uint64_t)n + 1, 0);
  for (uint32_t v : h.pins) h.ioff[v + 1]++;
  for (uint32_t v = 0; v < n; ++v) h.ioff[v + 1] += h.ioff[v];
  h.inc.resize(h.pins.size());
  {
    std::vector<uint64_t> at(h.ioff.begin(), h.ioff.end() - 1);
    for (uint64_t x = 0; x + 1 < h.off.size(); ++x)
      for (uint64_t i = h.off[x]; i < h.off[x + 1]; ++i) h.inc[at[h.pins[i]]++] = (uint32_t)x;
  }
  try {
    hgp_ip::Options opt;
    if (g_ip_threads == 1) opt.parallel_depth = 0;  // every try and half on the calling thread
    hgp_ip::partition(h, k, eps, omega, part, psizes, opt);
  } catch (const hgp_ip::Infeasible&) {
    return 6;  // HGP_E_INFEASIBLE
  } catch (const std::exception&) {
    return 8;  // HGP_E_INTERNAL
  }
  return 0;
}

// Whole pipeline: the recursion of sources/main.cu:282-784 with its base cases
// (KWAY :358-391 and :448-485, INCC :495-564), uncoarsening and refinement on the way
// up, INCC small-partition merge and final zero-based relabel (:803-830).
int orc_partition(int mode, uint32_t omega, uint32_t delta, uint32_t max_parts, uint32_t kway, float eps, uint32_t cnc,
                  uint32_t refine_repeats, int ipm, int prune, int ip_device, const uint32_t* ip_options,
                  uint32_t lp_rounds, uint32_t n0, uint32_t E, const uint64_t* hoff0,
                  const uint32_t* pins0, const uint32_t* srcc0, const float* w, uint32_t* part_out,
                  uint32_t* num_levels) {
  const bool kw = mode == 1;
  if (kw && max_parts != kway) return 1;  // k-way sets max_parts = k (sources/runconfig.cpp:190-192)
  const uint32_t target_parts = std::min(max_parts, (uint32_t)((n0 + omega - 1) / omega));
  struct Lv {
    uint32_t n;
    std::vector<uint64_t> hoff;
    std::vector<uint32_t> pins, srcc, sizes;
    std::vector<uint64_t> toff;
    std::vector<uint32_t> tch, inb;
    std::vector<uint32_t> groups;  // to the next level
  };
  std::vector<Lv> lv(1);
  lv[0].n = n0;
  lv[0].hoff.assign(hoff0, hoff0 + E + 1);
  lv[0].pins.assign(pins0, pins0 + hoff0[E]);
  lv[0].srcc.assign(srcc0, srcc0 + E);
  lv[0].sizes.assign(n0, 1u);
  touching(n0, E, lv[0].hoff.data(), lv[0].pins.data(), lv[0].srcc.data(), lv[0].toff, lv[0].tch, lv[0].inb);
  std::vector<uint32_t> part, psizes, pinb;
  uint32_t ip_opt[4] = {2, 8, 16, 4};
  for (int i = 0; i < 4; ++i)
    if (ip_options && ip_options[i]) ip_opt[i] = ip_options[i];
  // the k-way base case: the device partitioner's restatement (with the product's default
  // options: 2 starts, 8 rounds, 16 nodes per part, 4 refinement rounds) or the bisection
  auto base_ip = [&](uint32_t n, const uint64_t* hoff, const uint32_t* pins, const uint32_t* sizes) -> int {
    if (ip_device == 1 || (ip_device == 2 && hoff[E] >= 2097152u)) {  // device, or auto on big levels
      device_ip(n, E, hoff, pins, w, sizes, kway, omega, (uint32_t)std::max(prune, 0), ip_opt[0], ip_opt[1],
                ip_opt[2], ip_opt[3], part.data(), psizes.data());
      return 0;
    }
    return orc_initial_partition(n, E, hoff, pins, w, sizes, kway, eps, omega, prune, part.data(), psizes.data());
  };
  uint32_t K = 0;
  bool base_refined_level0 = false;
  uint32_t depth = 0;
  for (;; ++depth) {
    Lv& L = lv[depth];
    const uint32_t n = L.n;
    if (kw && n < 8192) {  // KWAY base case before coarsening
      part.assign(n, 0);
      psizes.assign(kway, 0);
      if (int rc = base_ip(n, L.hoff.data(), L.pins.data(), L.sizes.data()))
        return rc;
      K = kway;
      break;
    }
    std::vector<uint32_t> pairs((uint64_t)n * cnc), scores((uint64_t)n * cnc);
    orc_candidates(n, E, L.hoff.data(), L.pins.data(), L.srcc.data(), L.toff.data(), L.tch.data(), L.inb.data(), w,
                   L.sizes.data(), cnc, omega, delta, pairs.data(), scores.data());
    uint32_t nn = 0;
    std::vector<uint32_t> groups(n), gs(n), ung(n);
    std::vector<uint64_t> uoff(n + 1);
    orc_group_nodes(n, L.inb.data(), pairs.data(), scores.data(), L.sizes.data(), cnc, omega, delta, &nn,
                    groups.data(), gs.data(), ung.data(), uoff.data(), nullptr);
    gs.resize(nn);
    float ratio = (float)nn / (float)n;
    if (kw && (ratio > 0.95f || nn < 1024)) {
      part.assign(n, 0);
      psizes.assign(kway, 0);
      if (int rc = base_ip(n, L.hoff.data(), L.pins.data(), L.sizes.data()))
        return rc;
      K = kway;
      break;
    }
    if (!kw && (nn <= target_parts || nn == n || (depth >= 32 && ratio > 0.95f))) {
      if (ipm) orc_greedy_merge_groups(n, nn, L.sizes.data(), L.inb.data(), omega, delta, groups.data(), gs.data(), &nn);
      gs.resize(nn);
      part = groups;
      psizes = gs;
      K = nn;
      break;
    }
    L.groups = groups;
    Lv next;
    next.n = nn;
    coarsen_hedges(E, L.hoff.data(), L.pins.data(), L.srcc.data(), groups.data(), next.hoff, next.pins, next.srcc);
    next.sizes = gs;
    touching(nn, E, next.hoff.data(), next.pins.data(), next.srcc.data(), next.toff, next.tch, next.inb);
    lv.push_back(std::move(next));
  }
  // uncoarsen + refine back up to level 0
  pinb.assign(K, 0);
  for (int d = (int)depth - 1; d >= 0; --d) {
    Lv& L = lv[d];
    std::vector<uint32_t> fine(L.n);
    for (uint32_t v = 0; v < L.n; ++v) fine[v] = part[L.groups[v]];
    part.swap(fine);
    Level LL;
    std::vector<float> wv(w, w + E);
    build_level(LL, L.n, E, L.hoff.data(), L.pins.data(), L.srcc.data(), L.toff.data(), L.tch.data(), L.inb.data(),
                wv.data(), L.sizes.data());
    if (kw) lp_refine(LL, omega, K, lp_rounds, part.data(), psizes.data());
    refine(LL, mode, omega, delta, refine_repeats, K, d == 0, part.data(), psizes.data(), pinb.data());
    if (d == 0) base_refined_level0 = true;
  }
  if (!kw) {
    if (!base_refined_level0) orc_inbound_sizes(E, hoff0, pins0, srcc0, part.data(), K, pinb.data());
    merge_small(n0, K, psizes.data(), pinb.data(), omega, delta, part.data());
  }
  relabel(n0, part.data());
  std::copy(part.begin(), part.end(), part_out);
  if (num_levels) *num_levels = depth + 1;
  return 0;
}

// getPartitionsHypergraph(part, 2, true) (includes/hgraph.hpp:286-376): per hedge the set of
// source partitions and the set of destination partitions minus the source ones; dropped
// when at most one source partition and no destination remain; equal (sources, destinations)
// merged with float weights summed in hedge order.  Exact set keys (the reference keys a map
// by a 64-bit hash of the sets) and merged hedges in order of first appearance.  Two-phase:
// with null arrays only the sizes are returned.  Returns 1 when the partition ids are not
// 0..k-1 with every id used (the reference throws).
int orc_quotient(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                 const float* w, const uint32_t* part, uint32_t* k_out, uint32_t* e_out, uint64_t* p_out,
                 uint64_t* qoff, uint32_t* qpins, uint32_t* qsrcc, float* qw) {
  uint32_t k = 0;
  for (uint32_t v = 0; v < n; ++v) k = std::max(k, part[v] + 1);
  std::vector<uint8_t> used(k, 0);
  for (uint32_t v = 0; v < n; ++v) used[part[v]] = 1;
  for (uint32_t q = 0; q < k; ++q)
    if (!used[q]) return 1;
  std::map<std::pair<std::vector<uint32_t>, std::vector<uint32_t>>, uint32_t> index;
  std::vector<std::pair<std::vector<uint32_t>, std::vector<uint32_t>>> sets;
  std::vector<float> ws;
  for (uint32_t x = 0; x < e; ++x) {
    std::set<uint32_t> sp, dp;
    for (uint64_t i = hoff[x]; i < hoff[x + 1]; ++i) {
      if (i - hoff[x] < srcc[x]) sp.insert(part[pins[i]]);
      else dp.insert(part[pins[i]]);
    }
    for (uint32_t s : sp) dp.erase(s);
    if (sp.size() <= 1 && dp.empty()) continue;
    std::pair<std::vector<uint32_t>, std::vector<uint32_t>> key{{sp.begin(), sp.end()}, {dp.begin(), dp.end()}};
    auto it = index.find(key);
    if (it != index.end()) {
      ws[it->second] += w[x];
      continue;
    }
    index.emplace(key, (uint32_t)sets.size());
    sets.push_back(key);
    ws.push_back(w[x]);
  }
  uint64_t total = 0;
  for (auto& s : sets) total += s.first.size() + s.second.size();
  *k_out = k;
  *e_out = (uint32_t)sets.size();
  *p_out = total;
  if (!qoff) return 0;
  qoff[0] = 0;
  uint64_t at = 0;
  for (size_t g = 0; g < sets.size(); ++g) {
    for (uint32_t s : sets[g].first) qpins[at++] = s;
    for (uint32_t d : sets[g].second) qpins[at++] = d;
    qoff[g + 1] = at;
    qsrcc[g] = (uint32_t)sets[g].first.size();
    qw[g] = ws[g];
  }
  return 0;
}

int orc_check_validity(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                       const uint32_t* part, uint32_t omega, uint32_t delta, uint32_t max_parts) {
  std::unordered_map<uint32_t, uint32_t> counter;
  for (uint32_t v = 0; v < n; ++v) counter[part[v]]++;
  uint32_t count = (uint32_t)counter.size();
  if (count > max_parts) return 0;
  for (auto& kv : counter)
    if (kv.second > omega) return 0;
  for (uint32_t i = 0; i < count; ++i)
    if (!counter.count(i)) return -1;
  std::vector<uint32_t> syn(count, 0);
  std::unordered_set<uint32_t> seen;
  for (uint32_t h = 0; h < e; ++h) {
    seen.clear();
    for (uint64_t i = hoff[h] + srcc[h]; i < hoff[h + 1]; ++i)
      if (seen.insert(part[pins[i]]).second) syn[part[pins[i]]]++;
  }
  for (uint32_t i = 0; i < count; ++i)
    if (syn[i] > delta) return 0;
  return 1;
}

int orc_metrics(uint32_t n, uint32_t e, const uint64_t* hoff, const uint32_t* pins, const uint32_t* srcc,
                const float* w, const uint32_t* part, double* km1, double* cutnet, double* soed) {
  (void)n;
  double a = 0, b = 0, c = 0;
  std::set<uint32_t> sp, dp, all;
  for (uint32_t h = 0; h < e; ++h) {
    sp.clear();
    dp.clear();
    all.clear();
    for (uint64_t i = hoff[h]; i < hoff[h + 1]; ++i) {
      uint32_t q = part[pins[i]];
      all.insert(q);
      if (i - hoff[h] < srcc[h]) sp.insert(q);
      else dp.insert(q);
    }
    for (uint32_t s : sp) dp.erase(s);
    size_t len = sp.size() + dp.size();
    if (!(sp.size() <= 1 && dp.empty())) {
      a += (double)w[h] * (double)(len - 1);
      if (len > 1) b += w[h];
    }
    c += (double)w[h] * (double)(all.size() - 1) * (double)(hoff[h + 1] - hoff[h]);
  }
  *km1 = a;
  *cutnet = b;
  *soed = c;
  return 0;
}

}  // extern "C"
