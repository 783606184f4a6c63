// CAVLC of one 4x4 or 2x2 block of quantised levels: the syntax of
// CAVLCH264Impl::Encode / Decode (src/CodecUtils/CAVLCH264Impl.cpp:134-577)
// as host/device code over a bit writer / reader, with the H.264 VLC tables
// (coeff_token 9-5, total_zeros 9-7/9-8 and 9-9a, run_before 9-10) written
// out as the standard prints them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icu {
namespace cavlc {

struct Vlc {
  uint8_t len;    // 0: no such symbol
  uint16_t code;  // right-aligned
};

// "0001 01" -> {6, 0b000101}
constexpr Vlc V(const char* s) {
  Vlc v{0, 0};
  for (; *s; s++)
    if (*s == '0' || *s == '1') {
      v.code = (uint16_t)((v.code << 1) | (*s - '0'));
      v.len++;
    }
  return v;
}
constexpr Vlc kNone{0, 0};

// coeff_token[table][TrailingOnes][TotalCoeff]; tables: 0 <= nC < 2,
// 2 <= nC < 4, 4 <= nC < 8, nC == -1, nC == -2 (8 <= nC is a 6-bit fixed
// length code, see coeff_token()).
constexpr Vlc kCoeffToken[5][4][17] = {
    {  // 0 <= nC < 2
     {V("1"), V("0001 01"), V("0000 0111"), V("0000 0011 1"), V("0000 0001 11"), V("0000 0000 111"),
      V("0000 0000 0111 1"), V("0000 0000 0101 1"), V("0000 0000 0100 0"), V("0000 0000 0011 11"),
      V("0000 0000 0010 11"), V("0000 0000 0001 111"), V("0000 0000 0001 011"), V("0000 0000 0000 1111"),
      V("0000 0000 0000 1011"), V("0000 0000 0000 0111"), V("0000 0000 0000 0100")},
     {kNone, V("01"), V("0001 00"), V("0000 0110"), V("0000 0011 0"), V("0000 0001 10"), V("0000 0000 110"),
      V("0000 0000 0111 0"), V("0000 0000 0101 0"), V("0000 0000 0011 10"), V("0000 0000 0010 10"),
      V("0000 0000 0001 110"), V("0000 0000 0001 010"), V("0000 0000 0000 001"), V("0000 0000 0000 1110"),
      V("0000 0000 0000 1010"), V("0000 0000 0000 0110")},
     {kNone, kNone, V("001"), V("0000 101"), V("0000 0101"), V("0000 0010 1"), V("0000 0001 01"),
      V("0000 0000 101"), V("0000 0000 0110 1"), V("0000 0000 0100 1"), V("0000 0000 0011 01"),
      V("0000 0000 0010 01"), V("0000 0000 0001 101"), V("0000 0000 0001 001"), V("0000 0000 0000 1101"),
      V("0000 0000 0000 1001"), V("0000 0000 0000 0101")},
     {kNone, kNone, kNone, V("0001 1"), V("0000 11"), V("0000 100"), V("0000 0100"), V("0000 0010 0"),
      V("0000 0001 00"), V("0000 0000 100"), V("0000 0000 0110 0"), V("0000 0000 0011 00"),
      V("0000 0000 0010 00"), V("0000 0000 0001 100"), V("0000 0000 0001 000"), V("0000 0000 0000 1100"),
      V("0000 0000 0000 1000")}},
    {  // 2 <= nC < 4
     {V("11"), V("0010 11"), V("0001 11"), V("0000 111"), V("0000 0111"), V("0000 0100"), V("0000 0011 1"),
      V("0000 0001 111"), V("0000 0001 011"), V("0000 0000 1111"), V("0000 0000 1011"), V("0000 0000 1000"),
      V("0000 0000 0111 1"), V("0000 0000 0101 1"), V("0000 0000 0011 1"), V("0000 0000 0010 01"),
      V("0000 0000 0001 11")},
     {kNone, V("10"), V("0011 1"), V("0010 10"), V("0001 10"), V("0000 110"), V("0000 0110"), V("0000 0011 0"),
      V("0000 0001 110"), V("0000 0001 010"), V("0000 0000 1110"), V("0000 0000 1010"), V("0000 0000 0111 0"),
      V("0000 0000 0101 0"), V("0000 0000 0010 11"), V("0000 0000 0010 00"), V("0000 0000 0001 10")},
     {kNone, kNone, V("011"), V("0010 01"), V("0001 01"), V("0000 101"), V("0000 0101"), V("0000 0010 1"),
      V("0000 0001 101"), V("0000 0001 001"), V("0000 0000 1101"), V("0000 0000 1001"), V("0000 0000 0110 1"),
      V("0000 0000 0100 1"), V("0000 0000 0011 0"), V("0000 0000 0010 10"), V("0000 0000 0001 01")},
     {kNone, kNone, kNone, V("0101"), V("0100"), V("0011 0"), V("0010 00"), V("0001 00"), V("0000 100"),
      V("0000 0010 0"), V("0000 0001 100"), V("0000 0001 000"), V("0000 0000 1100"), V("0000 0000 0110 0"),
      V("0000 0000 0100 0"), V("0000 0000 0000 1"), V("0000 0000 0001 00")}},
    {  // 4 <= nC < 8
     {V("1111"), V("0011 11"), V("0010 11"), V("0010 00"), V("0001 111"), V("0001 011"), V("0001 001"),
      V("0001 000"), V("0000 1111"), V("0000 1011"), V("0000 0111 1"), V("0000 0101 1"), V("0000 0100 0"),
      V("0000 0011 01"), V("0000 0010 01"), V("0000 0001 01"), V("0000 0000 01")},
     {kNone, V("1110"), V("0111 1"), V("0110 0"), V("0101 0"), V("0100 0"), V("0011 10"), V("0010 10"),
      V("0001 110"), V("0000 1110"), V("0000 1010"), V("0000 0111 0"), V("0000 0101 0"), V("0000 0011 1"),
      V("0000 0011 00"), V("0000 0010 00"), V("0000 0001 00")},
     {kNone, kNone, V("1101"), V("0111 0"), V("0101 1"), V("0100 1"), V("0011 01"), V("0010 01"), V("0001 101"),
      V("0001 010"), V("0000 1101"), V("0000 1001"), V("0000 0110 1"), V("0000 0100 1"), V("0000 0010 11"),
      V("0000 0001 11"), V("0000 0000 11")},
     {kNone, kNone, kNone, V("1100"), V("1011"), V("1010"), V("1001"), V("1000"), V("0110 1"), V("0011 00"),
      V("0001 100"), V("0000 1100"), V("0000 1000"), V("0000 0110 0"), V("0000 0010 10"), V("0000 0001 10"),
      V("0000 0000 10")}},
    {  // nC == -1 (4:2:0 chroma DC)
     {V("01"), V("0001 11"), V("0001 00"), V("0000 11"), V("0000 10"), kNone, kNone, kNone, kNone, kNone, kNone,
      kNone, kNone, kNone, kNone, kNone, kNone},
     {kNone, V("1"), V("0001 10"), V("0000 011"), V("0000 0011"), kNone, kNone, kNone, kNone, kNone, kNone,
      kNone, kNone, kNone, kNone, kNone, kNone},
     {kNone, kNone, V("001"), V("0000 010"), V("0000 0010"), kNone, kNone, kNone, kNone, kNone, kNone, kNone,
      kNone, kNone, kNone, kNone, kNone},
     {kNone, kNone, kNone, V("0001 01"), V("0000 000"), kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone,
      kNone, kNone, kNone, kNone}},
    {  // nC == -2 (4:2:2 chroma DC)
     {V("1"), V("0001 111"), V("0001 110"), V("0000 0011 1"), V("0000 0011 0"), V("0000 0001 11"),
      V("0000 0000 111"), V("0000 0000 0111"), V("0000 0000 0011 1"), kNone, kNone, kNone, kNone, kNone, kNone,
      kNone, kNone},
     {kNone, V("01"), V("0001 101"), V("0001 100"), V("0000 0010 1"), V("0000 0001 10"), V("0000 0000 110"),
      V("0000 0000 0110"), V("0000 0000 0101"), kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone},
     {kNone, kNone, V("001"), V("0001 011"), V("0001 010"), V("0000 0010 0"), V("0000 0001 01"),
      V("0000 0000 101"), V("0000 0000 0100"), kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone},
     {kNone, kNone, kNone, V("0000 1"), V("0000 01"), V("0001 001"), V("0001 000"), V("0000 0001 00"),
      V("0000 0000 100"), kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone}}};

// total_zeros of 4x4 blocks, [TotalCoeff - 1][total_zeros].
constexpr Vlc kTotalZeros4x4[15][16] = {
    {V("1"), V("011"), V("010"), V("0011"), V("0010"), V("0001 1"), V("0001 0"), V("0000 11"), V("0000 10"),
     V("0000 011"), V("0000 010"), V("0000 0011"), V("0000 0010"), V("0000 0001 1"), V("0000 0001 0"),
     V("0000 0000 1")},
    {V("111"), V("110"), V("101"), V("100"), V("011"), V("0101"), V("0100"), V("0011"), V("0010"), V("0001 1"),
     V("0001 0"), V("0000 11"), V("0000 10"), V("0000 01"), V("0000 00"), kNone},
    {V("0101"), V("111"), V("110"), V("101"), V("0100"), V("0011"), V("100"), V("011"), V("0010"), V("0001 1"),
     V("0001 0"), V("0000 01"), V("0000 1"), V("0000 00"), kNone, kNone},
    {V("0001 1"), V("111"), V("0101"), V("0100"), V("110"), V("101"), V("100"), V("0011"), V("011"), V("0010"),
     V("0001 0"), V("0000 1"), V("0000 0"), kNone, kNone, kNone},
    {V("0101"), V("0100"), V("0011"), V("111"), V("110"), V("101"), V("100"), V("011"), V("0010"), V("0000 1"),
     V("0001"), V("0000 0"), kNone, kNone, kNone, kNone},
    {V("0000 01"), V("0000 1"), V("111"), V("110"), V("101"), V("100"), V("011"), V("010"), V("0001"), V("001"),
     V("0000 00"), kNone, kNone, kNone, kNone, kNone},
    {V("0000 01"), V("0000 1"), V("101"), V("100"), V("011"), V("11"), V("010"), V("0001"), V("001"),
     V("0000 00"), kNone, kNone, kNone, kNone, kNone, kNone},
    {V("0000 01"), V("0001"), V("0000 1"), V("011"), V("11"), V("10"), V("010"), V("001"), V("0000 00"), kNone,
     kNone, kNone, kNone, kNone, kNone, kNone},
    {V("0000 01"), V("0000 00"), V("0001"), V("11"), V("10"), V("001"), V("01"), V("0000 1"), kNone, kNone, kNone,
     kNone, kNone, kNone, kNone, kNone},
    {V("0000 1"), V("0000 0"), V("001"), V("11"), V("10"), V("01"), V("0001"), kNone, kNone, kNone, kNone, kNone,
     kNone, kNone, kNone, kNone},
    {V("0000"), V("0001"), V("001"), V("010"), V("1"), V("011"), kNone, kNone, kNone, kNone, kNone, kNone, kNone,
     kNone, kNone, kNone},
    {V("0000"), V("0001"), V("01"), V("1"), V("001"), kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone,
     kNone, kNone, kNone},
    {V("000"), V("001"), V("1"), V("01"), kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone,
     kNone, kNone},
    {V("00"), V("01"), V("1"), kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone,
     kNone},
    {V("0"), V("1"), kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone,
     kNone}};

// total_zeros of 2x2 chroma DC blocks, [TotalCoeff - 1][total_zeros].
constexpr Vlc kTotalZeros2x2[3][4] = {{V("1"), V("01"), V("001"), V("000")},
                                      {V("1"), V("01"), V("00"), kNone},
                                      {V("1"), V("0"), kNone, kNone}};

// run_before, [min(zerosLeft, 7) - 1][run_before].
constexpr Vlc kRunBefore[7][15] = {
    {V("1"), V("0"), kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone},
    {V("1"), V("01"), V("00"), kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone},
    {V("11"), V("10"), V("01"), V("00"), kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone,
     kNone},
    {V("11"), V("10"), V("01"), V("001"), V("000"), kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone, kNone,
     kNone},
    {V("11"), V("10"), V("011"), V("010"), V("001"), V("000"), kNone, kNone, kNone, kNone, kNone, kNone, kNone,
     kNone, kNone},
    {V("11"), V("000"), V("001"), V("011"), V("010"), V("101"), V("100"), kNone, kNone, kNone, kNone, kNone, kNone,
     kNone, kNone},
    {V("111"), V("110"), V("101"), V("100"), V("011"), V("010"), V("001"), V("0001"), V("0000 1"), V("0000 01"),
     V("0000 001"), V("0000 0001"), V("0000 0000 1"), V("0000 0000 01"), V("0000 0000 001")}};

// Scan orders (raster index of each scan position).
constexpr uint8_t kZigZag4x4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
constexpr uint8_t kZigZag2x2[4] = {0, 1, 2, 3};

// The 8x8 zig-zag (frame) scan: anti-diagonals in turn, odd ones walked
// down-left (row rising), even ones up-right.
struct Scan64 {
  uint8_t v[64];
};
constexpr Scan64 make_zigzag8x8() {
  Scan64 z{};
  int k = 0;
  for (int d = 0; d < 15; d++)
    for (int j = 0; j < 8; j++) {
      const int row = (d % 2) ? j : 7 - j;
      const int col = d - row;
      if (col >= 0 && col < 8) z.v[k++] = (uint8_t)(row * 8 + col);
    }
  return z;
}
constexpr Scan64 kZigZag8x8 = make_zigzag8x8();
static_assert(kZigZag8x8.v[2] == 8 && kZigZag8x8.v[3] == 16 && kZigZag8x8.v[63] == 63, "8x8 zig-zag");

enum { kMode2x2 = 0, kMode4x4 = 1, kMode8x8 = 2 };
enum { kErrSymbol = -1, kErrStream = -2 };

__host__ __device__ inline const uint8_t* scan_order(int mode) {
  return mode == kMode2x2 ? kZigZag2x2 : (mode == kMode8x8 ? kZigZag8x8.v : kZigZag4x4);
}
__host__ __device__ inline int block_size(int mode) { return mode == kMode2x2 ? 4 : (mode == kMode8x8 ? 64 : 16); }

// coeff_token code for (TotalCoeff, TrailingOnes) under nC.
__host__ __device__ inline Vlc coeff_token(int tc, int t1, int nc) {
  if (nc >= 8) {
    if (tc == 0) return Vlc{6, 3};
    return Vlc{6, (uint16_t)(((tc - 1) << 2) | t1)};
  }
  const int t = nc < 0 ? (nc == -1 ? 3 : 4) : (nc < 2 ? 0 : (nc < 4 ? 1 : 2));
  return kCoeffToken[t][t1][tc];
}

// Writes bits MSB-first; Sink provides put(value, nbits) (nbits <= 32),
// which returns false when the destination is full, and fits(nbits).  A
// symbol is written whole or not at all; on an error the symbols written
// before it stay, as with the reference's writer.
//
// 8x8 blocks (CAVLCH264Impl::Mode8x8, src/CodecUtils/CAVLCH264Impl.cpp:588-611)
// scan 64 levels but keep the 4x4 tables: the reference's coder indexes its
// coeff_token table with TotalCoeff and its total_zeros table with
// (total_zeros, TotalCoeff) as they come (CoeffTokenH264VlcEncoder.cpp:85-126,
// TotalZeros4x4H264VlcEncoder.cpp:83-92).  Within those tables (at most 15
// levels, total_zeros at most 15) the code is the 4x4 one over the 8x8
// scan, and a (total_zeros, TotalCoeff) pair the table has no code for fails
// after coeff_token and the levels are written, as there.  16 levels or more,
// or 16 zeros and more below the last level, index past the reference's
// tables; such a block is refused here before anything is written.
template <class Sink>
__host__ __device__ inline int encode_block(const int16_t* coef, int mode, int nc, int dc_skip, Sink& out) {
  const uint8_t* zz = scan_order(mode);
  const int maxc = block_size(mode);
  // levels in reverse scan order and the zero runs between them
  int level[16], run[16];
  int tc = 0, t1 = 0, tz = 0, zeros = 0;
  for (int i = maxc - 1; i >= dc_skip; i--) {
    const int x = coef[zz[i]];
    if (x == 0) {
      if (tc > 0) {
        zeros++;
        tz++;
      }
      continue;
    }
    if (tc == 16) return kErrSymbol;  // a 17th level: 8x8 blocks only
    if (tc > 0) run[tc - 1] = zeros;
    zeros = 0;
    level[tc++] = x;
    // trailing ones: +-1 levels at the high-frequency end, at most 3
    if (t1 == tc - 1 && t1 < 3 && (x == 1 || x == -1)) t1++;
  }
  if (mode == kMode8x8 && (tc > 15 || tz > 15)) return kErrSymbol;
  int bits = 0;
  const Vlc ct = coeff_token(tc, t1, nc);
  if (ct.len == 0) return kErrSymbol;
  if (!out.put(ct.code, ct.len)) return kErrStream;
  bits += ct.len;
  if (tc == 0) return bits;
  int sl = (tc > 10 && t1 < 3) ? 1 : 0;  // suffixLength
  for (int i = 0; i < tc; i++) {
    if (i < t1) {
      if (!out.put(level[i] < 0 ? 1u : 0u, 1)) return kErrStream;
      bits++;
      continue;
    }
    int lc = level[i] > 0 ? 2 * level[i] - 2 : -2 * level[i] - 1;  // levelCode
    if (i == t1 && t1 < 3) lc -= 2;
    int prefix, ssize, suffix;
    if (sl == 0 && lc >= 14 && lc <= 29) {
      prefix = 14;  // escape of the first suffix-less range
      ssize = 4;
      suffix = lc - 14;
    } else if (lc < (15 << sl)) {
      prefix = lc >> sl;
      ssize = sl;
      suffix = lc & ((1 << sl) - 1);
    } else {
      lc -= 15 << sl;
      if (sl == 0) lc -= 15;
      prefix = 15;
      while (prefix < 35 && lc >= (1LL << (prefix - 2)) - 4096) prefix++;
      if (prefix >= 16) lc -= (int)((1LL << (prefix - 3)) - 4096);
      ssize = prefix - 3;
      suffix = ssize >= 32 ? lc : (lc & (int)((1u << ssize) - 1u));
    }
    // level_prefix (prefix zeros then a one) and the suffix go in together
    if (!out.fits(prefix + 1 + ssize)) return kErrStream;
    for (int z = prefix; z > 0; z -= 32)
      if (!out.put(0u, z > 32 ? 32 : z)) return kErrStream;
    if (!out.put(1u, 1)) return kErrStream;
    if (ssize > 0 && !out.put((uint32_t)suffix, ssize)) return kErrStream;
    bits += prefix + 1 + ssize;
    if (sl == 0) sl = 1;
    const int a = level[i] < 0 ? -level[i] : level[i];
    if (a > (3 << (sl - 1)) && sl < 6) sl++;
  }
  if (tc < maxc - dc_skip) {
    const Vlc t = mode == kMode2x2 ? (tc <= 3 ? kTotalZeros2x2[tc - 1][tz] : kNone) : kTotalZeros4x4[tc - 1][tz];
    if (t.len == 0) return kErrSymbol;
    if (!out.put(t.code, t.len)) return kErrStream;
    bits += t.len;
    int left = tz;
    for (int i = 0; i < tc - 1 && left > 0; i++) {
      const Vlc r = kRunBefore[(left > 7 ? 7 : left) - 1][run[i]];
      if (r.len == 0) return kErrSymbol;
      if (!out.put(r.code, r.len)) return kErrStream;
      bits += r.len;
      left -= run[i];
    }
  }
  return bits;
}

// Source provides peek(n) (next n <= 32 bits, MSB-first, zeros past the
// end), skip(n) and left() (bits remaining).
template <class Source>
__host__ __device__ inline bool read_vlc(Source& in, const Vlc* tab, int n, int* sym, int* bits) {
  const uint32_t p = in.peek(16);
  for (int s = 0; s < n; s++) {
    const Vlc v = tab[s];
    if (v.len && (p >> (16 - v.len)) == v.code && v.len <= in.left()) {
      in.skip(v.len);
      *sym = s;
      *bits += v.len;
      return true;
    }
  }
  return false;
}

template <class Source>
__host__ __device__ inline int decode_block(Source& in, int mode, int nc, int dc_skip, int16_t* coef) {
  const uint8_t* zz = scan_order(mode);
  const int maxc = block_size(mode);
  for (int i = 0; i < maxc; i++) coef[i] = 0;
  if (in.left() <= 0) return kErrStream;
  int bits = 0, tc = -1, t1 = 0;
  // coeff_token
  const uint32_t p = in.peek(16);
  for (int t = 0; t < 4 && tc < 0; t++)
    for (int c = 0; c <= 16; c++) {
      const Vlc v = coeff_token(c, t, nc);
      if (v.len && (p >> (16 - v.len)) == v.code && v.len <= in.left()) {
        tc = c;
        t1 = t;
        in.skip(v.len);
        bits += v.len;
        break;
      }
    }
  if (tc < 0 || tc > maxc - dc_skip) return kErrSymbol;
  // 8x8: 16 levels would read total_zeros past the 4x4 table (see encode_block)
  if (mode == kMode8x8 && tc > 15) return kErrSymbol;
  if (tc == 0) return bits;
  int level[16], run[16];
  int sl = (tc > 10 && t1 < 3) ? 1 : 0;
  for (int i = 0; i < tc; i++) {
    if (i < t1) {
      if (in.left() <= 0) return kErrStream;
      level[i] = 1 - 2 * (int)in.peek(1);
      in.skip(1);
      bits++;
      continue;
    }
    int prefix = 0;
    while (true) {
      if (in.left() <= 0) return kErrStream;
      const uint32_t b = in.peek(1);
      in.skip(1);
      bits++;
      if (b) break;
      if (++prefix > 35) return kErrSymbol;
    }
    int ssize = sl;
    if (prefix >= 15) ssize = prefix - 3;
    else if (prefix == 14 && sl == 0) ssize = 4;
    int suffix = 0;
    if (ssize > 0) {
      if (in.left() < ssize) return kErrStream;
      suffix = (int)in.peek(ssize);
      in.skip(ssize);
      bits += ssize;
    }
    int lc = ((prefix < 15 ? prefix : 15) << sl) + suffix;
    if (prefix >= 15 && sl == 0) lc += 15;
    if (prefix >= 16) lc += (int)((1LL << (prefix - 3)) - 4096);
    if (i == t1 && t1 < 3) lc += 2;
    level[i] = (lc & 1) ? (-lc - 1) >> 1 : (lc + 2) >> 1;
    if (sl == 0) sl = 1;
    const int a = level[i] < 0 ? -level[i] : level[i];
    if (a > (3 << (sl - 1)) && sl < 6) sl++;
  }
  int left = 0;
  if (tc < maxc - dc_skip) {
    int tz;
    const bool ok = mode == kMode2x2 ? (tc <= 3 && read_vlc(in, kTotalZeros2x2[tc - 1], 4, &tz, &bits))
                                     : read_vlc(in, kTotalZeros4x4[tc - 1], 16, &tz, &bits);
    if (!ok) return in.left() <= 0 ? kErrStream : kErrSymbol;
    left = tz;
  }
  for (int i = 0; i < tc - 1; i++) {
    run[i] = 0;
    if (left > 0) {
      int r;
      if (!read_vlc(in, kRunBefore[(left > 7 ? 7 : left) - 1], 15, &r, &bits))
        return in.left() <= 0 ? kErrStream : kErrSymbol;
      if (r > left) return kErrSymbol;
      run[i] = r;
    }
    left -= run[i];
  }
  run[tc - 1] = left;
  int pos = -1;
  for (int i = tc - 1; i >= 0; i--) {
    pos += run[i] + 1;
    if (pos + dc_skip >= maxc) return kErrSymbol;
    coef[zz[pos + dc_skip]] = (int16_t)level[i];
  }
  return bits;
}

}  // namespace cavlc
}  // namespace icu
