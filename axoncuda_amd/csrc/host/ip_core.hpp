// ip_core.hpp - built-in k-way initial partitioner for the coarsest level (host, deterministic).
//
// The reference runs Mt-KaHyPar as an external process on the coarse hypergraph
// (sources/init_part.cu:733-953: drop hedges with fewer than two pins, merge identical
// hedges summing their 2^18 fixed-point weights, vertex weights = node sizes, -k k -e eps).
// When no `mtkahypar` binary is available the pipeline uses this partitioner instead:
// recursive bisection, each bisection = greedy hypergraph growing from several seeds
// followed by boundary FM passes with rollback, keeping the best cut.  Everything is
// integer and tie-broken by node id, so the result is a pure function of the input.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace hgp_ip {

// Raised when the size limit cannot hold the nodes in k partitions.
struct Infeasible : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Hypergraph {
  uint32_t n = 0;
  std::vector<uint64_t> off;   // hedge offsets
  std::vector<uint32_t> pins;  // sorted distinct pins per hedge
  std::vector<uint64_t> w;     // merged fixed-point weights
  std::vector<uint64_t> vw;    // node weights
  std::vector<uint64_t> ioff;  // incidence offsets
  std::vector<uint32_t> inc;   // incident hedges
};

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Parallel for over [0, m) in `threads` contiguous chunks (the calling thread takes one).
template <typename F>
inline void parallel_chunks(uint64_t m, uint32_t threads, F&& f) {
  if (threads <= 1 || m < 2) {
    f(0u, (uint64_t)0, m);
    return;
  }
  std::vector<std::thread> th;
  for (uint32_t t = 1; t < threads; ++t) th.emplace_back([&, t] { f(t, m * t / threads, m * (t + 1) / threads); });
  f(0u, (uint64_t)0, m / threads);
  for (auto& x : th) x.join();
}

// Incidence lists (hedges of each node, ascending).  Large inputs are split over threads by
// hedge range: per-thread node counts give every thread its own write positions, so the
// lists come out exactly as a sequential pass writes them.
inline void build_incidence(Hypergraph& h, uint32_t threads = 1) {
  const uint32_t m = (uint32_t)h.off.size() - 1;
  if ((uint64_t)h.n * threads > h.pins.size()) threads = 1;  // per-thread counts would dominate
  threads = std::max<uint32_t>(1, std::min<uint32_t>(threads, (uint32_t)(h.pins.size() >> 20) + 1));
  std::vector<uint64_t> cnt((uint64_t)threads * (h.n + 1), 0);
  parallel_chunks(m, threads, [&](uint32_t t, uint64_t lo, uint64_t hi) {
    uint64_t* c = cnt.data() + (uint64_t)t * (h.n + 1);
    for (uint64_t i = h.off[lo]; i < h.off[hi]; ++i) c[h.pins[i]]++;
  });
  h.ioff.assign(h.n + 1, 0);
  uint64_t run = 0;
  for (uint32_t v = 0; v < h.n; ++v) {
    h.ioff[v] = run;
    for (uint32_t t = 0; t < threads; ++t) {
      uint64_t& c = cnt[(uint64_t)t * (h.n + 1) + v];
      const uint64_t x = c;
      c = run;  // becomes thread t's write position for node v
      run += x;
    }
  }
  h.ioff[h.n] = run;
  h.inc.resize(h.pins.size());
  parallel_chunks(m, threads, [&](uint32_t t, uint64_t lo, uint64_t hi) {
    uint64_t* pos = cnt.data() + (uint64_t)t * (h.n + 1);
    for (uint64_t e = lo; e < hi; ++e)
      for (uint64_t i = h.off[e]; i < h.off[e + 1]; ++i) h.inc[pos[h.pins[i]]++] = (uint32_t)e;
  });
}

// Indexed binary max-heap keyed by (gain, -id).
class Heap {
 public:
  void reset(uint32_t n) {
    pos_.assign(n, UINT32_MAX);
    heap_.clear();
    key_.assign(n, 0);
  }
  bool contains(uint32_t v) const { return pos_[v] != UINT32_MAX; }
  bool empty() const { return heap_.empty(); }
  uint32_t top() const { return heap_[0]; }
  int64_t key(uint32_t v) const { return key_[v]; }
  void push(uint32_t v, int64_t k) {
    key_[v] = k;
    pos_[v] = (uint32_t)heap_.size();
    heap_.push_back(v);
    up(pos_[v]);
  }
  void update(uint32_t v, int64_t k) {
    int64_t old = key_[v];
    key_[v] = k;
    if (better(k, v, old, v)) up(pos_[v]); else down(pos_[v]);
  }
  void remove(uint32_t v) {
    uint32_t i = pos_[v];
    uint32_t last = heap_.back();
    heap_.pop_back();
    pos_[v] = UINT32_MAX;
    if (i < heap_.size()) {
      heap_[i] = last;
      pos_[last] = i;
      up(i);
      down(pos_[last]);
    }
  }

 private:
  static bool better(int64_t ka, uint32_t a, int64_t kb, uint32_t b) { return ka > kb || (ka == kb && a < b); }
  void up(uint32_t i) {
    while (i > 0) {
      uint32_t p = (i - 1) / 2;
      uint32_t a = heap_[i], b = heap_[p];
      if (!better(key_[a], a, key_[b], b)) break;
      heap_[i] = b; heap_[p] = a; pos_[b] = i; pos_[a] = p; i = p;
    }
  }
  void down(uint32_t i) {
    uint32_t n = (uint32_t)heap_.size();
    for (;;) {
      uint32_t l = 2 * i + 1, r = l + 1, m = i;
      if (l < n && better(key_[heap_[l]], heap_[l], key_[heap_[m]], heap_[m])) m = l;
      if (r < n && better(key_[heap_[r]], heap_[r], key_[heap_[m]], heap_[m])) m = r;
      if (m == i) break;
      uint32_t a = heap_[i], b = heap_[m];
      heap_[i] = b; heap_[m] = a; pos_[b] = i; pos_[a] = m; i = m;
    }
  }
  std::vector<uint32_t> heap_;
  std::vector<uint32_t> pos_;
  std::vector<int64_t> key_;
};

// The same queue as a dense array scanned for its maximum (key desc, id asc, as the heap):
// an update is O(1) and the maximum is recomputed only when the cached one is removed or
// lowered.  For the few thousand nodes of a coarsest level this beats the heap's sift per
// gain update by a wide margin.
class DenseQueue {
 public:
  void reset(uint32_t n) {
    key_.assign(n, kAbsent);
    count_ = 0;
    cached_ = UINT32_MAX;
  }
  bool contains(uint32_t v) const { return key_[v] != kAbsent; }
  bool empty() const { return count_ == 0; }
  uint32_t top() const {
    if (cached_ == UINT32_MAX) {
      int64_t best = kAbsent;
      uint32_t bi = UINT32_MAX;
      const int64_t* k = key_.data();
      for (uint32_t i = 0, n = (uint32_t)key_.size(); i < n; ++i)
        if (k[i] > best) {
          best = k[i];
          bi = i;
        }
      cached_ = bi;
    }
    return cached_;
  }
  // Best entry satisfying `ok` (one scan); UINT32_MAX when there is none.
  template <typename Ok>
  uint32_t best_if(Ok&& ok) const {
    int64_t best = kAbsent;
    uint32_t bi = UINT32_MAX;
    const int64_t* k = key_.data();
    for (uint32_t i = 0, n = (uint32_t)key_.size(); i < n; ++i)
      if (k[i] > best && ok(i)) {
        best = k[i];
        bi = i;
      }
    return bi;
  }
  void push(uint32_t v, int64_t k) {
    key_[v] = k;
    count_++;
    if (cached_ != UINT32_MAX && better(k, v, key_[cached_], cached_)) cached_ = v;
  }
  void update(uint32_t v, int64_t k) {
    const int64_t old = key_[v];
    key_[v] = k;
    if (cached_ == UINT32_MAX) return;
    if (v == cached_) {
      if (k < old) cached_ = UINT32_MAX;
    } else if (better(k, v, key_[cached_], cached_)) {
      cached_ = v;
    }
  }
  void remove(uint32_t v) {
    key_[v] = kAbsent;
    count_--;
    if (v == cached_) cached_ = UINT32_MAX;
  }

 private:
  static constexpr int64_t kAbsent = INT64_MIN;
  static bool better(int64_t ka, uint32_t a, int64_t kb, uint32_t b) { return ka > kb || (ka == kb && a < b); }
  std::vector<int64_t> key_;
  uint32_t count_ = 0;
  mutable uint32_t cached_ = UINT32_MAX;
};

// Gain queue of a bisection: dense scans for small sub-hypergraphs, a heap otherwise.
class GainQueue {
 public:
  void reset(uint32_t n) {
    dense_ = n <= kDenseMax;
    if (dense_) d_.reset(n);
    else h_.reset(n);
  }
  bool contains(uint32_t v) const { return dense_ ? d_.contains(v) : h_.contains(v); }
  bool empty() const { return dense_ ? d_.empty() : h_.empty(); }
  uint32_t top() const { return dense_ ? d_.top() : h_.top(); }
  void push(uint32_t v, int64_t k) { dense_ ? d_.push(v, k) : h_.push(v, k); }
  void update(uint32_t v, int64_t k) { dense_ ? d_.update(v, k) : h_.update(v, k); }
  void remove(uint32_t v) { dense_ ? d_.remove(v) : h_.remove(v); }
  // Best queued node satisfying `ok`.  The heap pops the failing tops on the way (into
  // `parked` when given; the caller re-queues them once they may fit again), which selects
  // the same node as the dense scan.
  template <typename Ok>
  uint32_t best_if(Ok&& ok, std::vector<uint32_t>* parked) {
    if (dense_) return d_.best_if(ok);
    while (!h_.empty()) {
      const uint32_t t = h_.top();
      if (ok(t)) return t;
      h_.remove(t);
      if (parked) parked->push_back(t);
    }
    return UINT32_MAX;
  }

 private:
  static constexpr uint32_t kDenseMax = 32768;
  bool dense_ = true;
  DenseQueue d_;
  Heap h_;
};

// Two-way partitioning of a sub-hypergraph (local ids 0..n-1).
// U = uint32_t packs a hedge into 16 bytes (pin offsets and weights below 2^32), uint64_t
// into 24 bytes for anything larger.
template <typename U>
struct BisectorT {
  // Everything a move reads or writes about one hedge, in one record (one cache line access
  // per incident hedge instead of one per array).  The pin count is c[0] + c[1].
  struct HRec {
    U off;          // first pin
    U w;            // weight
    uint32_t c[2];  // pins on side 0 / side 1
  };
  const Hypergraph& h;
  std::vector<HRec> rec;
  std::vector<uint8_t> side;
  std::vector<int64_t> gain;
  GainQueue heap[2];
  uint64_t wside[2] = {0, 0};
  bool gains_exact = false;  // gain[] holds every node's current move gain

  explicit BisectorT(const Hypergraph& hg) : h(hg) {
    const uint32_t m = (uint32_t)h.off.size() - 1;
    rec.resize(m);
    for (uint32_t e = 0; e < m; ++e) {
      rec[e].off = (U)h.off[e];
      rec[e].w = (U)h.w[e];
    }
  }
  static uint32_t len(const HRec& r) { return r.c[0] + r.c[1]; }

  int64_t cut() const {
    int64_t c = 0;
    for (const HRec& r : rec)
      if (r.c[0] && r.c[1]) c += (int64_t)r.w;
    return c;
  }
  void init_counts() {
    wside[0] = wside[1] = 0;
    for (uint32_t v = 0; v < h.n; ++v) wside[side[v]] += h.vw[v];
    for (uint32_t e = 0; e < (uint32_t)rec.size(); ++e) {
      HRec& r = rec[e];
      r.c[0] = r.c[1] = 0;
      for (uint64_t i = h.off[e]; i < h.off[e + 1]; ++i) r.c[side[h.pins[i]]]++;
    }
  }
  int64_t node_gain(uint32_t v) const {
    uint8_t a = side[v], b = 1 - a;
    int64_t g = 0;
    const uint64_t i0 = h.ioff[v], i1 = h.ioff[v + 1];
    for (uint64_t i = i0; i < i1; ++i) {
      if (i + 16 < i1) __builtin_prefetch(&rec[h.inc[i + 16]]);
      const HRec& r = rec[h.inc[i]];
      if (r.c[a] == 1) g += (int64_t)r.w;
      if (r.c[b] == 0) g -= (int64_t)r.w;
    }
    return g;
  }
  // Move v to the other side, updating the gains of free pins (classic FM delta rules,
  // all four applied in one scan of the hedge).
  // `frontier` (growing only): a hedge that gets its first side-b pin queues its side-a pins
  // that are not queued yet, with their final gains (later bumps update the queue).
  // `boundary` (FM): a hedge that this move cuts makes its free pins boundary nodes, queued
  // on their side when they are not queued already.
  void move(uint32_t v, const std::vector<uint8_t>& locked, std::vector<uint8_t>* frontier = nullptr,
            bool boundary = false) {
    const uint8_t a = side[v], b = 1 - a;
    const uint64_t i0 = h.ioff[v], i1 = h.ioff[v + 1];
    for (uint64_t i = i0; i < i1; ++i) {
      // the incident hedges are scattered over the record array: fetch ahead to overlap misses
      if (i + 16 < i1) __builtin_prefetch(&rec[h.inc[i + 16]], 1);
      if (i + 8 < i1) __builtin_prefetch(h.pins.data() + rec[h.inc[i + 8]].off);
      HRec& r = rec[h.inc[i]];
      const int64_t we = (int64_t)r.w;
      const uint32_t bprev = r.c[b], anew = r.c[a] - 1;
      r.c[a] = anew;
      r.c[b] = bprev + 1;
      // gain change of a free pin u on side a (da) and on side b (db)
      int64_t da = 0, db = 0;
      if (bprev == 0) da += we;    // the hedge stops being internal to side a
      if (bprev == 1) db -= we;    // the lone b pin no longer frees the hedge by moving
      if (anew == 0) db -= we;     // the hedge becomes internal to side b
      if (anew == 1) da += we;     // the lone a pin now frees the hedge by moving
      const bool queue = frontier && bprev == 0;
      const bool cut_now = boundary && bprev == 0 && anew >= 1;
      if (da == 0 && db == 0 && !queue && !cut_now) continue;
      const uint32_t* pp = h.pins.data() + r.off;
      const uint32_t rl = len(r);
      for (uint32_t j = 0; j < rl; ++j) {
        const uint32_t u = pp[j];
        if (u == v || locked[u]) continue;
        const int64_t d = side[u] == a ? da : db;
        if (d) bump(u, d);
        if (queue && side[u] == a && !(*frontier)[u]) {
          (*frontier)[u] = 1;
          heap[a].push(u, gain[u]);
        }
        if (cut_now && !heap[side[u]].contains(u)) heap[side[u]].push(u, gain[u]);
      }
    }
    side[v] = b;
    gain[v] = -gain[v];  // moving back undoes the move
    wside[a] -= h.vw[v];
    wside[b] += h.vw[v];
  }
  void bump(uint32_t u, int64_t d) {
    gain[u] += d;
    uint8_t s = side[u];
    if (heap[s].contains(u)) heap[s].update(u, gain[u]);
  }

  // Greedy growing: side 1 holds everything, side 0 grows from `seed` until it reaches target0.
  void grow(uint32_t seed, uint64_t target0, uint64_t max0) {
    // Everything starts on side 1: every hedge has all its pins there, and moving a node
    // to side 0 would cut every hedge it is on, so gain(v) = -(sum of its hedge weights).
    side.assign(h.n, 1);
    wside[0] = 0;
    wside[1] = 0;
    for (uint32_t v = 0; v < h.n; ++v) wside[1] += h.vw[v];
    gain.assign(h.n, 0);
    for (uint32_t e = 0; e < (uint32_t)rec.size(); ++e) {
      HRec& r = rec[e];
      r.c[0] = 0;
      r.c[1] = (uint32_t)(h.off[e + 1] - h.off[e]);
      for (uint64_t j = h.off[e]; j < h.off[e + 1]; ++j) gain[h.pins[j]] -= (int64_t)r.w;
    }
    std::vector<uint8_t> locked(h.n, 0);
    heap[0].reset(h.n);
    heap[1].reset(h.n);
    // Frontier: a hedge's pins are queued when its first pin moves to side 0 (inside move).
    std::vector<uint8_t> inq(h.n, 0);
    uint32_t next_unvisited = 0;
    uint32_t cur = seed;
    while (wside[0] < target0) {
      if (cur == UINT32_MAX) {
        // pick the best frontier node that fits (a node that does not fit never will: side 0
        // only grows); fall back to the next unvisited node
        cur = heap[1].best_if([&](uint32_t t) { return wside[0] + h.vw[t] <= max0; }, nullptr);
        if (cur == UINT32_MAX) {
          while (next_unvisited < h.n && (side[next_unvisited] == 0 || inq[next_unvisited] ||
                                          wside[0] + h.vw[next_unvisited] > max0))
            next_unvisited++;
          if (next_unvisited >= h.n) break;
          cur = next_unvisited;
        }
      }
      if (heap[1].contains(cur)) heap[1].remove(cur);
      inq[cur] = 1;
      move(cur, locked, &inq);
      cur = UINT32_MAX;
    }
    heap[1].reset(h.n);
    gains_exact = true;  // nothing was locked, so every gain was maintained
  }

  // One FM pass with rollback to the best prefix.  Returns the cut improvement.
  int64_t fm_pass(const uint64_t maxw[2]) {
    std::vector<uint8_t> locked(h.n, 0);
    heap[0].reset(h.n);
    heap[1].reset(h.n);
    for (uint32_t v = 0; v < h.n; ++v) {
      if (!gains_exact) gain[v] = node_gain(v);
      // boundary nodes only
      bool boundary = false;
      for (uint64_t i = h.ioff[v]; i < h.ioff[v + 1] && !boundary; ++i) {
        const HRec& r = rec[h.inc[i]];
        boundary = r.c[0] && r.c[1];
      }
      if (boundary) heap[side[v]].push(v, gain[v]);
    }
    gains_exact = false;  // moves below skip locked nodes and the rollback maintains nothing
    std::vector<uint32_t> moved, parked[2];
    int64_t cum = 0, best = 0;
    size_t best_len = 0;
    uint64_t best_imb = std::max(wside[0] > maxw[0] ? wside[0] - maxw[0] : 0,
                                 wside[1] > maxw[1] ? wside[1] - maxw[1] : 0);
    uint32_t stall = 0;
    const uint32_t stall_limit = std::max<uint32_t>(50, h.n / 32);
    while (!heap[0].empty() || !heap[1].empty()) {
      // candidate from each side that respects the target's capacity
      int pick = -1;
      uint32_t cand[2];
      for (int s = 0; s < 2; ++s)
        cand[s] = heap[s].best_if(
            [&](uint32_t t) { return wside[1 - s] + h.vw[t] <= maxw[1 - s] || wside[s] > maxw[s]; }, &parked[s]);
      int64_t g0 = cand[0] == UINT32_MAX ? INT64_MIN : gain[cand[0]];
      int64_t g1 = cand[1] == UINT32_MAX ? INT64_MIN : gain[cand[1]];
      if (g0 == INT64_MIN && g1 == INT64_MIN) break;
      if (g0 > g1) pick = 0;
      else if (g1 > g0) pick = 1;
      else pick = wside[0] >= wside[1] ? 0 : 1;
      uint32_t v = cand[pick];
      heap[pick].remove(v);
      locked[v] = 1;
      cum += gain[v];
      // hedges that this move cuts queue their free pins as boundary nodes (inside move)
      const uint8_t to = 1 - side[v];
      move(v, locked, nullptr, true);
      moved.push_back(v);
      // side `from` became lighter: nodes parked on side `to` for lack of room may fit now
      for (uint32_t u : parked[to])
        if (!locked[u] && side[u] == to && !heap[to].contains(u)) heap[to].push(u, gain[u]);
      parked[to].clear();
      uint64_t imb = std::max(wside[0] > maxw[0] ? wside[0] - maxw[0] : 0,
                              wside[1] > maxw[1] ? wside[1] - maxw[1] : 0);
      if (imb < best_imb || (imb == best_imb && cum > best)) {
        best = cum;
        best_len = moved.size();
        best_imb = imb;
        stall = 0;
      } else if (++stall > stall_limit) {
        break;
      }
    }
    std::vector<uint8_t> none(h.n, 1);  // no gain maintenance during rollback
    for (size_t i = moved.size(); i > best_len; --i) move(moved[i - 1], none);
    return best;
  }
};

// Restrict h to `nodes` (ascending ids): hedges keep their pins inside the subset, in
// hedge order, and those left with fewer than two pins are dropped.  Local ids follow the
// global order, so the pins of every hedge stay sorted.  Two sweeps over contiguous hedge
// ranges on `threads` threads (count, then fill at the range's own offsets), so the result
// is the same for any thread count.
inline Hypergraph subgraph(const Hypergraph& h, const std::vector<uint32_t>& nodes, uint32_t threads = 1) {
  Hypergraph s;
  s.n = (uint32_t)nodes.size();
  std::vector<uint32_t> local(h.n, UINT32_MAX);
  for (uint32_t i = 0; i < s.n; ++i) local[nodes[i]] = i;
  s.vw.resize(s.n);
  for (uint32_t i = 0; i < s.n; ++i) s.vw[i] = h.vw[nodes[i]];
  const uint32_t m = (uint32_t)h.off.size() - 1;
  threads = std::max<uint32_t>(1, std::min<uint32_t>(threads, (uint32_t)(h.pins.size() >> 18) + 1));
  std::vector<uint64_t> ch(threads + 1, 0), cp(threads + 1, 0);
  parallel_chunks(m, threads, [&](uint32_t t, uint64_t lo, uint64_t hi) {
    uint64_t nh = 0, np = 0;
    for (uint64_t e = lo; e < hi; ++e) {
      uint32_t c = 0;
      for (uint64_t j = h.off[e]; j < h.off[e + 1]; ++j) c += local[h.pins[j]] != UINT32_MAX;
      if (c >= 2) {
        nh++;
        np += c;
      }
    }
    ch[t + 1] = nh;
    cp[t + 1] = np;
  });
  for (uint32_t t = 0; t < threads; ++t) {
    ch[t + 1] += ch[t];
    cp[t + 1] += cp[t];
  }
  s.off.resize(ch[threads] + 1);
  s.w.resize(ch[threads]);
  s.pins.resize(cp[threads]);
  s.off[0] = 0;
  parallel_chunks(m, threads, [&](uint32_t t, uint64_t lo, uint64_t hi) {
    uint64_t k = ch[t], q = cp[t];
    for (uint64_t e = lo; e < hi; ++e) {
      uint32_t c = 0;
      for (uint64_t j = h.off[e]; j < h.off[e + 1]; ++j) c += local[h.pins[j]] != UINT32_MAX;
      if (c < 2) continue;  // writes stay inside the range counted for this chunk
      for (uint64_t j = h.off[e]; j < h.off[e + 1]; ++j) {
        const uint32_t x = local[h.pins[j]];
        if (x != UINT32_MAX) s.pins[q++] = x;
      }
      s.w[k] = h.w[e];
      s.off[++k] = q;
    }
  });
  build_incidence(s, threads);
  return s;
}

#ifndef HGP_IP_TRIES
#define HGP_IP_TRIES 4
#endif
#ifndef HGP_IP_FM_PASSES
#define HGP_IP_FM_PASSES 1
#endif

struct Options {
  uint32_t tries = HGP_IP_TRIES;          // growing seeds per bisection
  uint32_t fm_passes = HGP_IP_FM_PASSES;  // FM passes after each growing
  uint64_t seed = 86;
  uint32_t parallel_depth = 4;         // threads for tries/halves above this recursion depth
  uint64_t parallel_min_pins = 20000;  // smaller subproblems stay on the calling thread
};

struct TryResult {
  int64_t cut = INT64_MAX;
  uint64_t imb = UINT64_MAX;
  std::vector<uint8_t> side;
};

// One bisection attempt: greedy growing from a seed, then FM passes.
inline void bisect_try(const Hypergraph& h, uint32_t seed, uint64_t target0, const uint64_t maxw[2],
                       const Options& opt, TryResult& out) {
  auto run = [&](auto& b) {
    b.grow(seed, target0, maxw[0]);
    // FM passes until one gains nothing or less than 1/1000 of the cut (after the second)
    for (uint32_t p = 0; p < opt.fm_passes; ++p) {
      const int64_t g = b.fm_pass(maxw);
      if (p > 0 && (g <= 0 || g < b.cut() / 1000)) break;
    }
    out.cut = b.cut();
    out.imb = std::max(b.wside[0] > maxw[0] ? b.wside[0] - maxw[0] : 0, b.wside[1] > maxw[1] ? b.wside[1] - maxw[1] : 0);
    out.side = std::move(b.side);
  };
  bool narrow = h.pins.size() < (1ull << 32);
  for (uint64_t x : h.w) narrow = narrow && x < (1ull << 32);
  if (narrow) {
    BisectorT<uint32_t> b(h);
    run(b);
  } else {
    BisectorT<uint64_t> b(h);
    run(b);
  }
}

// Recursive bisection.  The tries of one bisection run on their own threads and the two
// halves recurse concurrently (down to depth opt.parallel_depth); every thread owns its
// state and the best try is chosen in try order, so the result does not depend on timing.
inline void bisect_recursive(const Hypergraph& h, const std::vector<uint32_t>& gids, uint32_t k, uint32_t first,
                             double eps_b, uint64_t omega, const Options& opt, uint32_t* part, uint32_t depth) {
  if (k == 1 || h.n == 0) {
    for (uint32_t g : gids) part[g] = first;
    return;
  }
  uint32_t k0 = k / 2, k1 = k - k0;
  uint64_t total = 0;
  for (uint64_t x : h.vw) total += x;
  uint64_t target0 = (uint64_t)((long double)total * k0 / k);
  uint64_t maxw[2];
  maxw[0] = std::min<uint64_t>((uint64_t)omega * k0, (uint64_t)((long double)total * k0 / k * (1.0 + eps_b)));
  maxw[1] = std::min<uint64_t>((uint64_t)omega * k1, (uint64_t)((long double)total * k1 / k * (1.0 + eps_b)));
  if (maxw[0] + maxw[1] < total) {  // keep the two capacities jointly feasible
    maxw[0] = std::min<uint64_t>((uint64_t)omega * k0, total);
    maxw[1] = std::min<uint64_t>((uint64_t)omega * k1, total);
  }
  std::vector<TryResult> res(opt.tries);
  std::vector<uint32_t> seeds(opt.tries);
  for (uint32_t t = 0; t < opt.tries; ++t)
    seeds[t] = (uint32_t)(splitmix64(opt.seed * 1315423911ull + depth * 2654435761ull + first * 97ull + t) % h.n);
  const bool par = depth < opt.parallel_depth && (uint64_t)h.pins.size() >= opt.parallel_min_pins;
  if (par) {
    std::vector<std::thread> th;
    for (uint32_t t = 1; t < opt.tries; ++t)
      th.emplace_back(bisect_try, std::cref(h), seeds[t], target0, maxw, std::cref(opt), std::ref(res[t]));
    if (opt.tries) bisect_try(h, seeds[0], target0, maxw, opt, res[0]);
    for (auto& x : th) x.join();
  } else {
    for (uint32_t t = 0; t < opt.tries; ++t) bisect_try(h, seeds[t], target0, maxw, opt, res[t]);
  }
  uint32_t bi = 0;
  for (uint32_t t = 1; t < opt.tries; ++t)
    if (res[t].imb < res[bi].imb || (res[t].imb == res[bi].imb && res[t].cut < res[bi].cut)) bi = t;
  const std::vector<uint8_t>& best_side = res[bi].side;
  std::vector<uint32_t> l[2], g[2];
  for (uint32_t v = 0; v < h.n; ++v) {
    l[best_side[v]].push_back(v);
    g[best_side[v]].push_back(gids[v]);
  }
  res.clear();
  // threads for building a half: the cores left to each of the 2^(depth+1) concurrent halves
  const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint32_t sub_threads = par ? std::max<uint32_t>(1, hw >> (depth + 1)) : 1u;
  auto half = [&](int s) {
    Hypergraph sub = subgraph(h, l[s], sub_threads);
    bisect_recursive(sub, g[s], s == 0 ? k0 : k1, s == 0 ? first : first + k0, eps_b, omega, opt, part, depth + 1);
  };
  if (par) {
    std::thread t0(half, 0);
    half(1);
    t0.join();
  } else {
    half(0);
    half(1);
  }
}

// Moves nodes out of overweight parts into the lightest parts that can take them
// (largest-first, ties by id) so that every part respects omega when possible.
inline void repair_balance(const Hypergraph& h, uint32_t k, uint64_t omega, uint32_t* part) {
  std::vector<uint64_t> pw(k, 0);
  for (uint32_t v = 0; v < h.n; ++v) pw[part[v]] += h.vw[v];
  for (uint32_t p = 0; p < k; ++p) {
    if (pw[p] <= omega) continue;
    std::vector<uint32_t> members;
    for (uint32_t v = 0; v < h.n; ++v)
      if (part[v] == p) members.push_back(v);
    std::sort(members.begin(), members.end(), [&](uint32_t a, uint32_t b) {
      return h.vw[a] != h.vw[b] ? h.vw[a] < h.vw[b] : a < b;
    });
    for (uint32_t v : members) {
      if (pw[p] <= omega) break;
      uint32_t dest = UINT32_MAX;
      for (uint32_t q = 0; q < k; ++q)
        if (q != p && pw[q] + h.vw[v] <= omega && (dest == UINT32_MAX || pw[q] < pw[dest])) dest = q;
      if (dest == UINT32_MAX) continue;
      pw[p] -= h.vw[v];
      pw[dest] += h.vw[v];
      part[v] = dest;
    }
  }
}

// Entry point: part[n] in [0,k), psizes[k] = summed node weights.
inline void partition(const Hypergraph& h, uint32_t k, float eps, uint64_t omega, uint32_t* part, uint32_t* psizes,
                      const Options& opt = Options()) {
  // Constraints that no k-partition meets: a node heavier than the size limit, or more
  // weight than k parts of that limit hold.  Checked before any part is assigned.
  uint64_t total = 0, heaviest = 0;
  for (uint32_t v = 0; v < h.n; ++v) {
    total += h.vw[v];
    heaviest = std::max<uint64_t>(heaviest, h.vw[v]);
  }
  if (k == 0 || heaviest > omega || total > (uint64_t)k * omega)
    throw Infeasible("initial partitioning: " + std::to_string(h.n) + " nodes of total weight " +
                     std::to_string(total) + " (heaviest " + std::to_string(heaviest) + ") do not fit " +
                     std::to_string(k) + " partitions of size limit " + std::to_string(omega));
  std::vector<uint32_t> gids(h.n);
  for (uint32_t v = 0; v < h.n; ++v) gids[v] = v;
  uint32_t levels = 0;
  while ((1u << levels) < k) levels++;
  double eps_b = levels ? std::pow(1.0 + (double)eps, 1.0 / levels) - 1.0 : (double)eps;
  if (k <= 1) {
    for (uint32_t v = 0; v < h.n; ++v) part[v] = 0;
  } else {
    bisect_recursive(h, gids, k, 0, eps_b, omega, opt, part, 0);
    repair_balance(h, k, omega, part);
  }
  for (uint32_t p = 0; p < k; ++p) psizes[p] = 0;
  for (uint32_t v = 0; v < h.n; ++v) psizes[part[v]] += (uint32_t)h.vw[v];
}

}  // namespace hgp_ip
