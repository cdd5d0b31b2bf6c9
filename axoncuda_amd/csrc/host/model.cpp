// model.cpp - host loaders, savers, metrics, validity and seeded generators.
//
// Loaders follow the reference formats and rules (includes/hgraph.hpp): .hgr (:936-1013;
// first non-comment line "E V [fmt]", 1-based pins, fmt 1/11 weights, vertex weights read
// but not used, single-pin hedges skipped), .snn (:707-743), .axh (:805-854).  Files are
// read whole and parsed without iostreams so that multi-GB inputs load at disk speed.
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <unordered_map>
#include <unordered_set>

namespace hgp {

namespace {

bool read_file(const std::string& path, std::vector<char>& buf) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  buf.resize((size_t)std::max<long>(sz, 0) + 1);
  size_t got = sz > 0 ? std::fread(buf.data(), 1, (size_t)sz, f) : 0;
  std::fclose(f);
  buf.resize(got + 1);
  buf[got] = '\0';
  return true;
}

bool ends_with(const std::string& s, const char* suf) {
  size_t l = std::strlen(suf);
  return s.size() >= l && s.compare(s.size() - l, l, suf) == 0;
}

struct LineCursor {
  const char* p;
  const char* end;
  // next data line (comments start with '%' after leading blanks); returns [b, e)
  bool next(const char*& b, const char*& e) {
    while (p < end) {
      const char* ls = p;
      const char* le = (const char*)std::memchr(p, '\n', (size_t)(end - p));
      if (!le) le = end;
      p = le < end ? le + 1 : end;
      const char* q = ls;
      while (q < le && (*q == ' ' || *q == '\t')) ++q;
      const char* qe = le;
      if (qe > q && qe[-1] == '\r') --qe;
      if (q >= qe) continue;
      if (*q == '%') continue;
      b = q;
      e = qe;
      return true;
    }
    return false;
  }
};

inline bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// parses an unsigned integer token; returns false when the next token is not a number
inline bool parse_uint(const char*& q, const char* e, uint64_t& out, bool& negative) {
  while (q < e && is_blank(*q)) ++q;
  negative = false;
  if (q >= e) return false;
  if (*q == '-' || *q == '+') {
    negative = *q == '-';
    ++q;
  }
  if (q >= e || *q < '0' || *q > '9') return false;
  uint64_t v = 0;
  while (q < e && *q >= '0' && *q <= '9') {
    v = v * 10 + (uint64_t)(*q - '0');
    if (v > 0xFFFFFFFFFFull) v = 0xFFFFFFFFFFull;
    ++q;
  }
  out = v;
  return true;
}

// Host threads for the loaders and graders: at most 16, the process's share of a GPU box.
unsigned host_threads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

// fn(t, lo, hi) on T contiguous ranges of [0, count), one thread each (the caller's thread
// takes range 0).
template <typename F>
void parallel_ranges(uint64_t count, unsigned T, F&& fn) {
  if (T <= 1 || count < 2) {
    fn(0u, (uint64_t)0, count);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned t = 1; t < T; ++t) th.emplace_back([&, t] { fn(t, count * t / T, count * (t + 1) / T); });
  fn(0u, (uint64_t)0, count / T);
  for (auto& x : th) x.join();
}

// Parse of one hyperedge line [b, e): its pins (0-based) appended to `pins` and its weight;
// returns "" or the error text of the line.
const char* parse_hedge_line(const char* b, const char* e, bool ew, uint64_t V, std::vector<uint32_t>& pins,
                             float& w, size_t& cnt) {
  const char* q = b;
  bool neg;
  w = 1.0f;
  if (ew) {
    while (q < e && is_blank(*q)) ++q;
    char* endp = nullptr;
    w = std::strtof(q, &endp);
    if (endp == q) w = 1.0f;
    q = endp > e ? e : endp;
  }
  const size_t start = pins.size();
  uint64_t v;
  while (parse_uint(q, e, v, neg)) {
    if (neg || v == 0 || v > V) return "Invalid node encountered";
    pins.push_back((uint32_t)(v - 1));
    if (q < e && !is_blank(*q)) break;  // token continues with non-digits: extraction stops
  }
  cnt = pins.size() - start;
  if (cnt < 1) return "Hyperedge must contain at least one node";
  return "";
}

// Files above this size are parsed by host_threads() threads: the body after the header is cut
// into ranges at line ends, each range's data lines are counted, and then every range parses
// its lines knowing their global index (hyperedge i, or vertex weight i - E).  The first
// error in file order is reported, as by a sequential read.
constexpr size_t HGR_PARALLEL_BYTES = 4u << 20;

std::string load_hgr(const std::string& path, hgp_hgraph& hg) {
  std::vector<char> buf;
  if (!read_file(path, buf)) return "Cannot open .hgr file";
  LineCursor lc{buf.data(), buf.data() + buf.size() - 1};
  const char *b, *e;
  if (!lc.next(b, e)) return "Empty .hgr file";
  uint64_t E = 0, V = 0, fmt = 0;
  bool neg;
  const char* q = b;
  if (!parse_uint(q, e, E, neg) || !parse_uint(q, e, V, neg)) return "Invalid .hgr header";
  if (E > 0xFFFFFFFFull) return "Hyperedge count exceeds uint32 range.";
  if (V > 0xFFFFFFFFull) return "Node count exceeds uint32 range.";
  if (!parse_uint(q, e, fmt, neg)) fmt = 0;
  const bool ew = fmt == 1 || fmt == 11, vw = fmt == 10 || fmt == 11;
  hg = hgp_hgraph();
  hg.n = (uint32_t)V;
  const unsigned T = host_threads();
  if (buf.size() >= HGR_PARALLEL_BYTES && T > 1) {
    const char* body = lc.p;
    const char* bend = lc.end;
    // range starts at line starts
    std::vector<const char*> cut(T + 1);
    cut[0] = body;
    cut[T] = bend;
    for (unsigned t = 1; t < T; ++t) {
      const char* x = body + (size_t)(bend - body) * t / T;
      if (x < cut[t - 1]) x = cut[t - 1];
      const char* nl = (const char*)std::memchr(x, '\n', (size_t)(bend - x));
      cut[t] = nl ? nl + 1 : bend;
    }
    std::vector<uint64_t> lines(T + 1, 0);
    parallel_ranges(T, T, [&](unsigned, uint64_t lo, uint64_t hi) {
      for (uint64_t t = lo; t < hi; ++t) {
        LineCursor c{cut[t], cut[t + 1]};
        const char *xb, *xe;
        uint64_t k = 0;
        while (c.next(xb, xe)) ++k;
        lines[t + 1] = k;
      }
    });
    for (unsigned t = 0; t < T; ++t) lines[t + 1] += lines[t];
    if (vw) hg.vweights.assign(hg.n, 1);
    struct Part {
      std::vector<uint32_t> pins;
      std::vector<uint64_t> ends;  // end offset (local) of each kept hyperedge
      std::vector<float> w;
      uint64_t err_line = UINT64_MAX;
      const char* err = "";
    };
    std::vector<Part> parts(T);
    parallel_ranges(T, T, [&](unsigned, uint64_t lo, uint64_t hi) {
      for (uint64_t t = lo; t < hi; ++t) {
        Part& P = parts[t];
        LineCursor c{cut[t], cut[t + 1]};
        const char *xb, *xe;
        uint64_t g = lines[t];
        if (g < E) P.pins.reserve((size_t)(cut[t + 1] - cut[t]) / 4);
        while (c.next(xb, xe)) {
          if (g < E) {
            float w;
            size_t cnt;
            const size_t start = P.pins.size();
            const char* err = parse_hedge_line(xb, xe, ew, V, P.pins, w, cnt);
            if (*err) {
              P.err_line = g;
              P.err = err;
              break;
            }
            if (cnt == 1) {
              P.pins.resize(start);  // singleton hedges are skipped
            } else {
              P.ends.push_back(P.pins.size());
              P.w.push_back(w);
            }
          } else if (vw && g - E < V) {
            const char* q = xb;
            uint64_t x;
            bool ng;
            if (parse_uint(q, xe, x, ng)) hg.vweights[g - E] = (uint32_t)x;
          } else {
            break;
          }
          ++g;
        }
      }
    });
    for (unsigned t = 0; t < T; ++t)
      if (parts[t].err_line != UINT64_MAX) return parts[t].err;  // ranges are in file order
    if (lines[T] < E) return "Unexpected end of file while reading hyperedges";
    uint64_t total = 0, edges = 0;
    for (auto& P : parts) {
      total += P.pins.size();
      edges += P.ends.size();
    }
    hg.pins.resize(total);
    hg.off.resize(edges + 1);
    hg.srcc.assign(edges, 0);
    hg.w.resize(edges);
    std::vector<uint64_t> pbase(T + 1, 0), ebase(T + 1, 0);
    for (unsigned t = 0; t < T; ++t) {
      pbase[t + 1] = pbase[t] + parts[t].pins.size();
      ebase[t + 1] = ebase[t] + parts[t].ends.size();
    }
    hg.off[0] = 0;
    parallel_ranges(T, T, [&](unsigned, uint64_t lo, uint64_t hi) {
      for (uint64_t t = lo; t < hi; ++t) {
        const Part& P = parts[t];
        std::memcpy(hg.pins.data() + pbase[t], P.pins.data(), P.pins.size() * 4);
        for (size_t i = 0; i < P.ends.size(); ++i) {
          hg.off[ebase[t] + i + 1] = pbase[t] + P.ends[i];
          hg.w[ebase[t] + i] = P.w[i];
        }
      }
    });
    return "";
  }
  hg.pins.reserve(buf.size() / 4);
  hg.srcc.reserve(E);
  hg.w.reserve(E);
  hg.off.reserve(E + 1);
  for (uint64_t i = 0; i < E; ++i) {
    if (!lc.next(b, e)) return "Unexpected end of file while reading hyperedges";
    float w;
    size_t cnt;
    const size_t start = hg.pins.size();
    const char* err = parse_hedge_line(b, e, ew, V, hg.pins, w, cnt);
    if (*err) return err;
    if (cnt == 1) {
      hg.pins.resize(start);  // singleton hedges are skipped
      continue;
    }
    hg.off.push_back(hg.pins.size());
    hg.srcc.push_back(0);
    hg.w.push_back(w);
  }
  if (vw) {
    hg.vweights.assign(hg.n, 1);
    for (uint32_t i = 0; i < hg.n; ++i) {
      if (!lc.next(b, e)) break;
      q = b;
      uint64_t x;
      if (parse_uint(q, e, x, neg)) hg.vweights[i] = (uint32_t)x;
    }
  }
  return "";
}

std::string load_bin(const std::string& path, bool axh, hgp_hgraph& hg) {
  std::vector<char> buf;
  if (!read_file(path, buf)) return axh ? "Cannot open .axh file" : "Cannot open file";
  const uint8_t* p = (const uint8_t*)buf.data();
  const uint8_t* end = p + buf.size() - 1;
  auto r32 = [&](uint32_t& v) -> bool {
    if (end - p < 4) return false;
    std::memcpy(&v, p, 4);
    p += 4;
    return true;
  };
  uint32_t N, E;
  if (!r32(N) || !r32(E)) return "Truncated header";
  hg = hgp_hgraph();
  hg.n = N;
  hg.srcc.reserve(E);
  hg.w.reserve(E);
  hg.off.reserve((size_t)E + 1);
  for (uint32_t i = 0; i < E; ++i) {
    uint32_t a, b;
    if (!r32(a) || !r32(b)) return "Truncated hyperedge";
    uint32_t npins, nsrc;
    if (axh) {
      npins = a;
      nsrc = b;
      if (nsrc == 0 || npins < nsrc) return "Invalid AXH hyperedge header";
      for (uint32_t k = 0; k < npins; ++k) {
        uint32_t v;
        if (!r32(v)) return "Truncated hyperedge";
        if (v >= N) return "Invalid node id in AXH file";
        hg.pins.push_back(v);
      }
    } else {
      npins = a + 1;
      nsrc = 1;
      if (b >= N) return "Invalid source node";
      hg.pins.push_back(b);
      for (uint32_t k = 0; k < a; ++k) {
        uint32_t v;
        if (!r32(v)) return "Truncated hyperedge";
        if (v >= N) return "Invalid destination node";
        hg.pins.push_back(v);
      }
    }
    if (npins == nsrc) return "Hyperedge must have at least 1 destination";
    float w;
    if (end - p < 4) return "Truncated weight";
    std::memcpy(&w, p, 4);
    p += 4;
    hg.off.push_back(hg.pins.size());
    hg.srcc.push_back(nsrc);
    hg.w.push_back(w);
  }
  return "";
}

void put_u32(std::vector<char>& out, uint32_t v) {
  char b[4];
  std::memcpy(b, &v, 4);
  out.insert(out.end(), b, b + 4);
}

}  // namespace

std::string load_hgraph(const std::string& path, hgp_hgraph& hg) {
  if (ends_with(path, ".hgr")) return load_hgr(path, hg);
  if (ends_with(path, ".snn")) return load_bin(path, false, hg);
  if (ends_with(path, ".axh")) return load_bin(path, true, hg);
  return "Failed to load hypergraph, unsupported file format (supported: '.hgr', '.snn', '.axh').";
}

namespace {
// Rule-2 dedup of hyperedges [h0, h1) into `out` (sources first, each pin's first occurrence
// kept in order): `ends` gets each hedge's end in `out`, `nsrc` its source count.  Hedges of
// up to 32 pins compare against the pins already kept; longer ones use `stamp` (one entry per
// node, allocated on first use).
void dedup_range(const hgp_hgraph& hg, uint32_t h0, uint32_t h1, std::vector<uint32_t>& out,
                 std::vector<uint64_t>& ends, std::vector<uint32_t>& nsrc) {
  std::vector<uint32_t> stamp;
  uint32_t gen = 0;
  out.reserve((size_t)(hg.off[h1] - hg.off[h0]));
  ends.reserve(h1 - h0);
  nsrc.reserve(h1 - h0);
  for (uint32_t i = h0; i < h1; ++i) {
    const uint64_t s = hg.off[i], t = hg.off[i + 1];
    const size_t ws = out.size();
    const uint32_t sc = hg.srcc[i];
    uint32_t ns = 0;
    if (t - s <= 32) {
      for (uint64_t k = s; k < t; ++k) {
        const uint32_t v = hg.pins[k];
        bool seen = false;
        for (size_t x = ws; x < out.size() && !seen; ++x) seen = out[x] == v;
        if (!seen) {
          out.push_back(v);
          if (k < s + sc) ++ns;
        }
      }
    } else {
      if (stamp.empty()) stamp.assign(hg.n, 0);
      if (++gen == 0) {
        std::fill(stamp.begin(), stamp.end(), 0);
        gen = 1;
      }
      for (uint64_t k = s; k < t; ++k) {
        const uint32_t v = hg.pins[k];
        if (stamp[v] != gen) {
          stamp[v] = gen;
          out.push_back(v);
          if (k < s + sc) ++ns;
        }
      }
    }
    ends.push_back(out.size());
    nsrc.push_back(ns);
  }
}
}  // namespace

// Rule 2 (includes/hgraph.hpp:518-581): within each hedge a pin repeated among the sources, or
// among the destinations, or appearing as both, keeps only its first occurrence (sources are
// stored first, so a pin that is a source stays one).  Hedge ranges are deduplicated by
// host_threads() threads and concatenated in order.
void dedup_rule2(hgp_hgraph& hg) {
  const uint32_t E = hg.e();
  const unsigned T = hg.pins.size() >= (1u << 20) ? host_threads() : 1u;
  std::vector<std::vector<uint32_t>> outs(T), nsrcs(T);
  std::vector<std::vector<uint64_t>> endss(T);
  std::vector<uint32_t> bounds(T + 1);
  for (unsigned t = 0; t <= T; ++t) bounds[t] = (uint32_t)((uint64_t)E * t / T);
  parallel_ranges(T, T, [&](unsigned, uint64_t lo, uint64_t hi) {
    for (uint64_t t = lo; t < hi; ++t) dedup_range(hg, bounds[t], bounds[t + 1], outs[t], endss[t], nsrcs[t]);
  });
  std::vector<uint64_t> base(T + 1, 0);
  for (unsigned t = 0; t < T; ++t) base[t + 1] = base[t] + outs[t].size();
  hg.pins.resize(base[T]);
  hg.off[0] = 0;
  parallel_ranges(T, T, [&](unsigned, uint64_t lo, uint64_t hi) {
    for (uint64_t t = lo; t < hi; ++t) {
      std::memcpy(hg.pins.data() + base[t], outs[t].data(), outs[t].size() * 4);
      for (uint32_t i = bounds[t]; i < bounds[t + 1]; ++i) {
        hg.off[i + 1] = base[t] + endss[t][i - bounds[t]];
        hg.srcc[i] = nsrcs[t][i - bounds[t]];
      }
    }
  });
}

std::string save_hgraph(const hgp_hgraph& hg, const std::string& path) {
  std::vector<char> out;
  if (ends_with(path, ".hgr")) {
    // savehMETIS (:1016-1051): weights as u32(w * 1024) when any differs from 1
    bool emit = false;
    for (float x : hg.w)
      if (x != 1.0f) {
        emit = true;
        break;
      }
    std::string s = std::to_string(hg.e()) + " " + std::to_string(hg.n) + (emit ? " 01\n" : "\n");
    out.insert(out.end(), s.begin(), s.end());
    char tmp[32];
    for (uint32_t i = 0; i < hg.e(); ++i) {
      if (emit) {
        int l = std::snprintf(tmp, sizeof tmp, "%u ", (uint32_t)(hg.w[i] * 1024.0f));
        out.insert(out.end(), tmp, tmp + l);
      }
      for (uint64_t k = hg.off[i]; k < hg.off[i + 1]; ++k) {
        int l = std::snprintf(tmp, sizeof tmp, k + 1 < hg.off[i + 1] ? "%u " : "%u", hg.pins[k] + 1);
        out.insert(out.end(), tmp, tmp + l);
      }
      out.push_back('\n');
    }
  } else if (ends_with(path, ".snn") || ends_with(path, ".axh")) {
    const bool axh = ends_with(path, ".axh");
    put_u32(out, hg.n);
    put_u32(out, hg.e());
    for (uint32_t i = 0; i < hg.e(); ++i) {
      const uint32_t len = (uint32_t)(hg.off[i + 1] - hg.off[i]);
      if (axh) {
        if (hg.srcc[i] == 0 || len < hg.srcc[i]) return "Invalid hyperedge encountered while saving AXH";
        put_u32(out, len);
        put_u32(out, hg.srcc[i]);
      } else {
        // saveSNN (:746-774): first pin is the source, the rest destinations
        put_u32(out, len - 1);
      }
      for (uint64_t k = hg.off[i]; k < hg.off[i + 1]; ++k) put_u32(out, hg.pins[k]);
      char b[4];
      std::memcpy(b, &hg.w[i], 4);
      out.insert(out.end(), b, b + 4);
    }
  } else {
    return "Failed to save partitioned hypergraph, unsupported file format (supported: '.hgr', '.snn', '.axh').";
  }
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return "Cannot open output file";
  std::fwrite(out.data(), 1, out.size(), f);
  std::fclose(f);
  return "";
}

hgp_hgraph partitions_hypergraph(const hgp_hgraph& hg, const std::vector<uint32_t>& part) {
  uint32_t k = 0;
  for (uint32_t p : part) k = std::max(k, p + 1);
  struct Acc {
    std::vector<uint32_t> s, d;
    float w = 0;
  };
  std::unordered_map<std::string, size_t> index;
  std::vector<Acc> accs;
  std::vector<uint32_t> sp, dp;
  std::string key;
  for (uint32_t i = 0; i < hg.e(); ++i) {
    sp.clear();
    dp.clear();
    for (uint64_t j = hg.off[i]; j < hg.off[i + 1]; ++j) {
      if (j - hg.off[i] < hg.srcc[i]) sp.push_back(part[hg.pins[j]]);
      else dp.push_back(part[hg.pins[j]]);
    }
    std::sort(sp.begin(), sp.end());
    sp.erase(std::unique(sp.begin(), sp.end()), sp.end());
    std::sort(dp.begin(), dp.end());
    dp.erase(std::unique(dp.begin(), dp.end()), dp.end());
    std::vector<uint32_t> d2;
    for (uint32_t x : dp)
      if (!std::binary_search(sp.begin(), sp.end(), x)) d2.push_back(x);
    if (sp.size() <= 1 && d2.empty()) continue;
    key.assign((const char*)sp.data(), sp.size() * 4);
    key.push_back('|');
    key.append((const char*)d2.data(), d2.size() * 4);
    auto it = index.find(key);
    if (it == index.end()) {
      index.emplace(key, accs.size());
      accs.push_back(Acc{sp, d2, hg.w[i]});
    } else {
      accs[it->second].w += hg.w[i];
    }
  }
  hgp_hgraph q;
  q.n = k;
  for (auto& a : accs) {
    q.pins.insert(q.pins.end(), a.s.begin(), a.s.end());
    q.pins.insert(q.pins.end(), a.d.begin(), a.d.end());
    q.off.push_back(q.pins.size());
    q.srcc.push_back((uint32_t)a.s.size());
    q.w.push_back(a.w);
  }
  return q;
}

// lambda-1 and cut-net of the quotient hypergraph, SOED on the input (grading/grade.cpp:300-308)
Metrics metrics(const hgp_hgraph& hg, const std::vector<uint32_t>& part) {
  Metrics m;
  const uint32_t E = hg.e();
  // Per hedge: lam = distinct partitions of all its pins, ns = distinct source partitions,
  // nd = distinct destination partitions that hold no source.  Counted by host_threads()
  // threads with per-thread stamps indexed by partition id (sorted copies when the ids are
  // too large to index); the sums then run in hedge order, as a sequential count would.
  uint32_t maxq = 0;
  for (uint32_t q : part) maxq = std::max(maxq, q);
  std::vector<uint32_t> lam(E), nsv(E), ndv(E);
  const bool stamped = (uint64_t)maxq < (1ull << 22);  // 16 MB of stamps per thread at most
  const unsigned T = hg.pins.size() >= (1u << 20) ? host_threads() : 1u;
  parallel_ranges(E, T, [&](unsigned, uint64_t lo, uint64_t hi) {
    std::vector<uint32_t> s_all, sp, dp, all;
    if (stamped) s_all.assign((size_t)maxq + 1, UINT32_MAX);
    for (uint64_t i = lo; i < hi; ++i) {
      const uint64_t s = hg.off[i], t = hg.off[i + 1], sc = s + hg.srcc[i];
      if (stamped) {
        uint32_t l = 0, ns = 0, nd = 0;
        const uint32_t g = (uint32_t)i;
        for (uint64_t j = s; j < sc; ++j) {  // sources first
          const uint32_t q = part[hg.pins[j]];
          if (s_all[q] != g) {
            s_all[q] = g;
            ++l;
            ++ns;
          }
        }
        for (uint64_t j = sc; j < t; ++j) {  // a destination partition seen first here holds no source
          const uint32_t q = part[hg.pins[j]];
          if (s_all[q] != g) {
            s_all[q] = g;
            ++l;
            ++nd;
          }
        }
        lam[i] = l;
        nsv[i] = ns;
        ndv[i] = nd;
      } else {
        sp.clear();
        dp.clear();
        all.clear();
        for (uint64_t j = s; j < t; ++j) {
          const uint32_t q = part[hg.pins[j]];
          all.push_back(q);
          (j < sc ? sp : dp).push_back(q);
        }
        std::sort(all.begin(), all.end());
        lam[i] = (uint32_t)(std::unique(all.begin(), all.end()) - all.begin());
        std::sort(sp.begin(), sp.end());
        sp.erase(std::unique(sp.begin(), sp.end()), sp.end());
        std::sort(dp.begin(), dp.end());
        dp.erase(std::unique(dp.begin(), dp.end()), dp.end());
        uint32_t nd = 0;
        for (uint32_t x : dp)
          if (!std::binary_search(sp.begin(), sp.end(), x)) ++nd;
        nsv[i] = (uint32_t)sp.size();
        ndv[i] = nd;
      }
    }
  });
  for (uint32_t i = 0; i < E; ++i) {
    const size_t len = (size_t)nsv[i] + ndv[i];
    if (!(nsv[i] <= 1 && ndv[i] == 0)) {
      m.connectivity += (double)hg.w[i] * (double)(len - 1);
      if (len > 1) m.cutnet += hg.w[i];
      m.quotient_pins += len;
    }
    m.soed += (double)hg.w[i] * (double)(lam[i] - 1) * (double)(hg.off[i + 1] - hg.off[i]);
  }
  return m;
}

int validity(const hgp_hgraph& hg, const std::vector<uint32_t>& part, uint32_t omega, uint32_t delta, uint32_t max_parts,
             std::string* why) {
  // Partition sizes in an array indexed by id when the ids are small enough.  The map of the
  // sequential check (built in node order) is built only to name a partition over the size
  // limit, so that the message is the one that check gives.
  uint32_t maxq = 0;
  for (uint32_t p : part) maxq = std::max(maxq, p);
  const bool arr = !part.empty() && (uint64_t)maxq < (1ull << 24);
  std::vector<uint32_t> cnt;
  std::unordered_map<uint32_t, uint32_t> counter;
  uint32_t count = 0;
  bool over = false;
  if (arr) {
    cnt.assign((size_t)maxq + 1, 0);
    for (uint32_t p : part) cnt[p]++;
    for (uint32_t c : cnt) {
      count += c != 0;
      over |= c > omega;
    }
  } else {
    for (uint32_t p : part) counter[p]++;
    count = (uint32_t)counter.size();
  }
  char buf[256];
  if (count > max_parts) {
    if (why) {
      std::snprintf(buf, sizeof buf, "more partitions (%u) than allowed (%u)", count, max_parts);
      *why = buf;
    }
    return 0;
  }
  if (arr && over)
    for (uint32_t p : part) counter[p]++;
  for (auto& kv : counter)
    if (kv.second > omega) {
      if (why) {
        std::snprintf(buf, sizeof buf, "more nodes per partition (%u) than allowed (%u)", kv.second, omega);
        *why = buf;
      }
      return 0;
    }
  for (uint32_t i = 0; i < count; ++i)
    if (arr ? (i > maxq || cnt[i] == 0) : !counter.count(i)) {
      if (why) *why = "Partitions must be incrementally indexed from 0 onward.";
      return -1;
    }
  // inbound hedges per partition (a hedge counts once for each partition among its
  // destinations): per-thread counts over hedge ranges, summed (integers, so exact)
  const unsigned T = hg.pins.size() >= (1u << 20) ? host_threads() : 1u;
  std::vector<std::vector<uint32_t>> syns(T);
  parallel_ranges(hg.e(), T, [&](unsigned t, uint64_t lo, uint64_t hi) {
    std::vector<uint32_t>& syn = syns[t];
    syn.assign(count, 0);
    std::vector<uint32_t> stamp(count, UINT32_MAX);
    for (uint64_t i = lo; i < hi; ++i)
      for (uint64_t j = hg.off[i] + hg.srcc[i]; j < hg.off[i + 1]; ++j) {
        const uint32_t q = part[hg.pins[j]];
        if (stamp[q] != (uint32_t)i) {
          stamp[q] = (uint32_t)i;
          syn[q]++;
        }
      }
  });
  std::vector<uint32_t>& syn = syns[0];
  for (unsigned t = 1; t < T; ++t)
    for (uint32_t q = 0; q < count; ++q) syn[q] += syns[t][q];
  for (uint32_t i = 0; i < count; ++i)
    if (syn[i] > delta) {
      if (why) {
        std::snprintf(buf, sizeof buf, "more inbound hedges per partition (%u) than allowed (%u)", syn[i], delta);
        *why = buf;
      }
      return 0;
    }
  return 1;
}

// ---- generators -----------------------------------------------------------------------------

void gen_random(uint32_t n, uint32_t e, uint32_t max_size, uint64_t seed, hgp_hgraph& hg) {
  std::mt19937_64 rng(seed);
  std::geometric_distribution<uint32_t> geo(0.5);  // failures before success: X-1 for X on {1,2,...}
  std::uniform_int_distribution<uint32_t> node(0, n - 1);
  hg = hgp_hgraph();
  hg.n = n;
  std::vector<uint32_t> cur;
  for (uint32_t i = 0; i < e; ++i) {
    uint32_t sz = 2 + 1 + geo(rng);
    if (sz > max_size) sz = max_size;
    if (sz > n) sz = n;
    cur.clear();
    while (cur.size() < sz) {
      uint32_t v = node(rng);
      if (std::find(cur.begin(), cur.end(), v) == cur.end()) cur.push_back(v);
    }
    hg.pins.insert(hg.pins.end(), cur.begin(), cur.end());
    hg.off.push_back(hg.pins.size());
    hg.srcc.push_back(0);
    hg.w.push_back(1.0f);
  }
}

void gen_netlist(uint32_t n, uint32_t e, double alpha, uint32_t max_size, uint32_t radius, double frac_local,
                 uint64_t seed, hgp_hgraph& hg) {
  std::mt19937_64 rng(seed);
  // discrete power law on [2, max_size]: inverse CDF table
  std::vector<double> cdf;
  double tot = 0;
  for (uint32_t s = 2; s <= max_size; ++s) {
    tot += std::pow((double)s, -alpha);
    cdf.push_back(tot);
  }
  for (double& x : cdf) x /= tot;
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const uint32_t side = (uint32_t)std::ceil(std::sqrt((double)n));
  std::uniform_int_distribution<uint32_t> node(0, n - 1);
  std::uniform_int_distribution<int> off(-(int)radius, (int)radius);
  hg = hgp_hgraph();
  hg.n = n;
  hg.vweights.resize(n);
  std::uniform_int_distribution<uint32_t> vw(1, 4);
  for (uint32_t v = 0; v < n; ++v) hg.vweights[v] = vw(rng);
  std::vector<uint32_t> cur;
  for (uint32_t i = 0; i < e; ++i) {
    uint32_t sz = 2 + (uint32_t)(std::lower_bound(cdf.begin(), cdf.end(), U(rng)) - cdf.begin());
    if (sz > n) sz = n;
    cur.clear();
    const uint32_t first = node(rng);
    cur.push_back(first);
    const int fx = (int)(first % side), fy = (int)(first / side);
    uint32_t guard = 0;
    while (cur.size() < sz && guard++ < 64 * sz + 1024) {
      uint32_t v;
      if (U(rng) < frac_local) {
        int dx, dy;
        do {
          dx = off(rng);
          dy = off(rng);
        } while (dx * dx + dy * dy > (int)(radius * radius));
        const int x = fx + dx, y = fy + dy;
        if (x < 0 || y < 0 || x >= (int)side || y >= (int)side) continue;
        const uint64_t id = (uint64_t)y * side + (uint64_t)x;
        if (id >= n) continue;
        v = (uint32_t)id;
      } else {
        v = node(rng);
      }
      if (std::find(cur.begin(), cur.end(), v) == cur.end()) cur.push_back(v);
    }
    while (cur.size() < sz) {  // dense neighbourhood exhausted: complete uniformly
      uint32_t v = node(rng);
      if (std::find(cur.begin(), cur.end(), v) == cur.end()) cur.push_back(v);
    }
    hg.pins.insert(hg.pins.end(), cur.begin(), cur.end());
    hg.off.push_back(hg.pins.size());
    hg.srcc.push_back(0);
    hg.w.push_back(1.0f);
  }
}

void gen_snn(uint32_t layers, uint32_t per_layer, uint32_t grid, uint32_t fanout, double sigma, uint64_t seed,
             hgp_hgraph& hg) {
  // One generator stream per layer (seeded from `seed` and the layer index), so the layers
  // are drawn on their own threads and the result does not depend on the thread count.
  hg = hgp_hgraph();
  hg.n = layers * per_layer;
  const uint32_t L = layers - 1;
  std::vector<std::vector<uint32_t>> lp(L);
  auto layer = [&](uint32_t l) {
    std::mt19937_64 rng(seed * 0x9E3779B97F4A7C15ull + l + 1);
    std::normal_distribution<double> N01(0.0, 1.0);
    std::vector<uint32_t>& out = lp[l];
    out.reserve((size_t)per_layer * (fanout + 1));
    std::vector<uint32_t> stamp(per_layer, UINT32_MAX);
    for (uint32_t i = 0; i < per_layer; ++i) {
      const uint32_t src = l * per_layer + i;
      const double cx = (double)(i % grid), cy = (double)(i / grid);
      out.push_back(src);
      uint32_t got = 0;
      uint64_t guard = 0;
      while (got < fanout) {
        int x = (int)std::lround(cx + sigma * N01(rng));
        int y = (int)std::lround(cy + sigma * N01(rng));
        if (++guard > 1000ull * fanout) {  // pathological: fall back to uniform cells
          x = (int)(rng() % grid);
          y = (int)(rng() % grid);
        }
        x = std::min(std::max(x, 0), (int)grid - 1);
        y = std::min(std::max(y, 0), (int)grid - 1);
        const uint64_t cell = (uint64_t)y * grid + (uint64_t)x;
        if (cell >= per_layer || stamp[cell] == i) continue;  // off the layer, or a repeat
        stamp[cell] = i;
        out.push_back((l + 1) * per_layer + (uint32_t)cell);
        ++got;
      }
    }
  };
  std::vector<std::thread> th;
  for (uint32_t l = 0; l < L; ++l) th.emplace_back(layer, l);
  for (auto& t : th) t.join();
  uint64_t total = 0;
  for (auto& v : lp) total += v.size();
  hg.pins.reserve(total);
  hg.off.reserve((size_t)L * per_layer + 1);
  for (uint32_t l = 0; l < L; ++l) {
    const uint64_t base = hg.pins.size();
    hg.pins.insert(hg.pins.end(), lp[l].begin(), lp[l].end());
    std::vector<uint32_t>().swap(lp[l]);
    for (uint32_t i = 0; i < per_layer; ++i) {
      hg.off.push_back(base + (uint64_t)(i + 1) * (fanout + 1));
      hg.srcc.push_back(1);
      hg.w.push_back(1.0f);
    }
  }
}

}  // namespace hgp
