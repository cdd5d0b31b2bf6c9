// prim.hpp - in-house device-wide primitives (replacing the reference's thrust/CUB calls:
// inclusive/exclusive scans, stable radix sorts, segment offsets, reductions).
#pragma once

#include "common.hpp"

namespace hgp {

// out[0..n) = exclusive prefix of in[0..n); out[n] = total.  In-place (out == in) is
// allowed when the element types match.
void exclusive_scan(Ctx& c, const uint32_t* in, uint64_t* out, uint64_t n);
void exclusive_scan(Ctx& c, const uint64_t* in, uint64_t* out, uint64_t n);
// out[i] = sum in[0..i]; in-place allowed.
void inclusive_scan(Ctx& c, const int64_t* in, int64_t* out, uint64_t n);
void inclusive_scan(Ctx& c, const int32_t* in, int32_t* out, uint64_t n);

// Stable LSD radix sorts over key bits [0, bits); values ride along (nullable).
void radix_sort(Ctx& c, uint32_t* keys, uint32_t* vals, uint64_t n, int bits);
void radix_sort(Ctx& c, uint64_t* keys, uint32_t* vals, uint64_t n, int bits);
// Stable sort by key bits [lo, hi) only: when the input is already ordered by the bits below
// lo, the result is ordered by bits [0, hi), in (hi - lo) / 8 passes instead of hi / 8.
void radix_sort_bits(Ctx& c, uint64_t* keys, uint32_t* vals, uint64_t n, int lo, int hi);
// The same over bits [lo, hi) of buffers the caller owns: after an odd number of passes the
// buffers are exchanged with the sort's scratch ones instead of copied back (the DBufs then
// hold different allocations; any raw pointer taken before is stale).
void radix_sort_buf(Ctx& c, DBuf<uint32_t>& keys, DBuf<uint32_t>& vals, uint64_t n, int lo, int hi);
void radix_sort_buf(Ctx& c, DBuf<uint64_t>& keys, DBuf<uint32_t>& vals, uint64_t n, int lo, int hi);
// Stable sort over the keys' actual range: digits are taken from key - min, so only the bits
// that the maximum difference needs are sorted; the keys come out unchanged.
void radix_sort_range(Ctx& c, uint64_t* keys, uint32_t* vals, uint64_t n);
// The same on buffers the caller owns, exchanged with the scratch ones instead of copied back
// (see radix_sort_buf).
void radix_sort_range_buf(Ctx& c, DBuf<uint64_t>& keys, DBuf<uint32_t>& vals, uint64_t n);

// off[s] = first index i with (keys[i] >> shift) >= s, for s in [0, nseg]; keys sorted.
void segment_offsets(Ctx& c, const uint32_t* keys, int shift, uint64_t n, uint32_t nseg, uint64_t* off);
void segment_offsets(Ctx& c, const uint64_t* keys, int shift, uint64_t n, uint32_t nseg, uint64_t* off);

void fill_u32(Ctx& c, uint32_t* p, uint32_t v, uint64_t n);
void iota_u32(Ctx& c, uint32_t* p, uint64_t n);
uint64_t reduce_max_u64(Ctx& c, const uint64_t* p, uint64_t n);
uint32_t reduce_max_u32(Ctx& c, const uint32_t* p, uint64_t n);

inline int bit_width(uint64_t x) {
  int b = 0;
  while (x) {
    b++;
    x >>= 1;
  }
  return b;
}

}  // namespace hgp
