// Per-macroblock decision, sub-pel refinement and final vector choice, as
// device functions shared by the decision kernels and the search kernel
// (which runs them on the MBs of a task right after its integer search).
//
// One wavefront (64 lanes) decides one MB.  It restates, in order, the rest
// of MotionEstimatorH264ImplFull::Estimate
// (src/CodecUtils/MotionEstimatorH264ImplFull.cpp:350-601):
//   * the integer decision, either from the search kernel's packed argmin
//     (zero predictor, filter off) or by replaying the reference's scan over
//     the cost table: with the running-best filter (:410-411) the scan is
//     path dependent, so lanes test 64 consecutive candidates at once against
//     the current best, a ballot finds the first one the reference would
//     accept, and the scan restarts after it;
//   * the half-pel stage (LoadHalfQuartPelWindow :1095-1182, loop :452-483)
//     from a 35x35 grid of G/b/h/j samples built in LDS;
//   * the quarter-pel stage (LoadQuartPelWindow :1198-1222, loop :485-517)
//     with the estimator's own pairwise averages;
//   * the predictor distortion (:529-570) and GetBestVector
//     (MotionEstimatorH264ImplFull.h:140-168) in correctly rounded doubles.
// With the median predictor (H264MotionVectorPredictorImpl1 over
// MacroBlockH264::GetMbMotionMedianPred, MacroBlockH264.cpp:757-816) an MB
// depends on its left, above-left, above and above-right neighbours, so MB
// rows run as a wavefront: each row is owned by one wave that polls the row
// above's vector words until they are final.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "icu_device.h"
#include "icu_internal.h"
#include "icu_qpel.h"

namespace icu {
namespace dec {


constexpr int kHalf = 35;       // half-pel grid side


struct WaveLds {
  uint32_t cur4[64];                // the MB's input block, 16 rows of 4 dwords
  uint8_t P[kPatch * kPatch];       // reference patch
  int16_t hv[17 * kPatch];          // raw vertical 6-tap sums
  int16_t Hg[kHalf * kHalf];        // G/b/h/j grid on the half-pel lattice
};

// Half-pel neighbour offsets in the reference's candidate order
// (Full.cpp:452-483): row above left to right, left, right, row below.
constexpr int kSubX[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
constexpr int kSubY[8] = {-1, -1, -1, 0, 0, 1, 1, 1};

// The MB's cost-table window [yu..yd] x [xl..xr], read straight from the
// table (each entry is visited about once, so it is not staged).
struct TabView {
  const uint32_t* t;
  int dim, L, xl, yu, w;
  __device__ __forceinline__ uint32_t operator[](int idx) const {
    const int i = idx / w, j = idx - i * w;
    return t[(i + yu + L) * dim + (j + xl + L)];
  }
};

// LDS written by one lane of a wave and read by another, later in program
// order: a wavefront-scope fence plus a wave barrier keeps the compiler from
// reordering across it (a wave's LDS operations complete in order).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Wave sums by DPP, within the VALU: four row shifts (1, 2, 4, 8; lanes
// shifted in from outside the 16-lane row read 0) leave each lane the sum
// of its row up to itself, so lane 16 r + 15 holds row r's sum; row_bcast:15
// then adds row 0's into row 1 and row 2's into row 3, and row_bcast:31 adds
// lane 31 into rows 2 and 3, which leaves the total in lane 63.  Every
// caller runs these with all 64 lanes active.
__device__ __forceinline__ int row_scan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);  // row_shr:8
  return v;
}
__device__ __forceinline__ int rows_to_total(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
  return __builtin_amdgcn_readlane(v, 63);
}
// The sum over the wave, the same in every lane.
__device__ __forceinline__ int wave_sum(int v) { return rows_to_total(row_scan(v)); }
// Eight wave sums at once: the DPP steps of one value are independent of
// the others', so they interleave.
__device__ __forceinline__ void wave_sum8(int (&a)[8]) {
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] += __builtin_amdgcn_update_dpp(0, a[k], 0x111, 0xf, 0xf, true);
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] += __builtin_amdgcn_update_dpp(0, a[k], 0x112, 0xf, 0xf, true);
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] += __builtin_amdgcn_update_dpp(0, a[k], 0x114, 0xf, 0xf, true);
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] += __builtin_amdgcn_update_dpp(0, a[k], 0x118, 0xf, 0xf, true);
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] += __builtin_amdgcn_update_dpp(0, a[k], 0x142, 0xa, 0xf, false);
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] += __builtin_amdgcn_update_dpp(0, a[k], 0x143, 0xc, 0xf, false);
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] = __builtin_amdgcn_readlane(a[k], 63);
}
// Eight wave sums whose totals are below 2^16 (a 16x16 SAD is at most
// 65280): packed in pairs into 32-bit words, no partial sum carries into
// its neighbour, so four sums of words do the work of eight.
__device__ __forceinline__ void wave_sum8_small(int (&a)[8]) {
  int w[4];
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] = a[2 * i] | (a[2 * i + 1] << 16);
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] += __builtin_amdgcn_update_dpp(0, w[i], 0x111, 0xf, 0xf, true);
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] += __builtin_amdgcn_update_dpp(0, w[i], 0x112, 0xf, 0xf, true);
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] += __builtin_amdgcn_update_dpp(0, w[i], 0x114, 0xf, 0xf, true);
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] += __builtin_amdgcn_update_dpp(0, w[i], 0x118, 0xf, 0xf, true);
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] += __builtin_amdgcn_update_dpp(0, w[i], 0x142, 0xa, 0xf, false);
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] += __builtin_amdgcn_update_dpp(0, w[i], 0x143, 0xc, 0xf, false);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint32_t t = (uint32_t)__builtin_amdgcn_readlane(w[i], 63);
    a[2 * i] = (int)(t & 0xFFFFu);
    a[2 * i + 1] = (int)(t >> 16);
  }
}
// The least 64-bit key over the wave as a uniform value: the least high
// word, then the least low word among the lanes that hold it.  Every caller
// runs it with all 64 lanes active.
__device__ __forceinline__ unsigned long long wave_min_u64_uniform(unsigned long long v) {
  const uint32_t hi = wave_min_u32_uniform((uint32_t)(v >> 32));
  const uint32_t lo = wave_min_u32_uniform((uint32_t)(v >> 32) == hi ? (uint32_t)v : 0xFFFFFFFFu);
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}

__device__ __forceinline__ int pix_dist(int a, int b, int metric) {
  const int d = a - b;
  return metric == kMetricSAD ? (d < 0 ? -d : d) : d * d;
}

// acc + the distance of pels a and b (0..255) under a metric known at
// compile time: for SAD one v_sad_u16 (|a - b| + acc on the low halves).
template <int METRIC>
__device__ __forceinline__ int dist_add(int acc, int a, int b) {
  if constexpr (METRIC == kMetricSAD) {
    return (int)__builtin_amdgcn_sad_u16((uint32_t)a, (uint32_t)b, (uint32_t)acc);
  } else {
    const int e = a - b;
    return acc + e * e;
  }
}

// Stage the 22x22 patch whose (3,3) is reference pel (x0, y0): 22 rows of 7
// dwords read from the dword boundary at or left of x0-3 (the padded plane's
// origin and stride are multiples of 16), every load issued before the first
// LDS store.
__device__ inline void load_patch(const DecideDesc& d, int x0, int y0, uint8_t* P, int lane) {
  const int xa = x0 - 3, sh = xa & 3;  // 0..3 also for negative xa
  const uint32_t* base = (const uint32_t*)(d.ref + (long)(y0 - 3) * d.ref_stride + (xa - sh));
  const int sdw = d.ref_stride >> 2;
  uint32_t w[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const int i = lane + 64 * j, r = i / 7, k = i - 7 * r;
    w[j] = i < 7 * kPatch ? base[r * sdw + k] : 0u;
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const int i = lane + 64 * j, r = i / 7, k = i - 7 * r;
    if (i < 7 * kPatch) {
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int c = 4 * k + b - sh;
        if (c >= 0 && c < kPatch) P[r * kPatch + c] = (uint8_t)(w[j] >> (8 * b));
      }
    }
  }
}

// Integer decision by argmin over the table with the predictor-distance key
// (filter off): lower cost, then closer to (pXr, pYr), then later in raster.
__device__ inline void table_argmin(const TabView& tab, int xl, int xr, int yu, int yd, int pXr, int pYr,
                             int lane, int* mx, int* my, int* minD) {
  const int w = xr - xl + 1, n = w * (yd - yu + 1);
  unsigned long long best = ~0ull;
  for (int idx = lane; idx < n; idx += 64) {
    const int i = yu + idx / w, j = xl + idx % w;
    const uint32_t c = tab[idx];
    unsigned long long rr, d2;
    if (i == 0 && j == 0) {
      rr = 0x1FFFFull;
    } else {
      rr = 0xFFFFull - (unsigned long long)((i + 64) * 256 + (j + 64));
    }
    const int ex = j - pXr, ey = i - pYr;
    d2 = (unsigned long long)(ex * ex + ey * ey);
    const unsigned long long key = ((unsigned long long)c << 40) | (d2 << 17) | rr;
    best = key < best ? key : best;
  }
  best = wave_min_u64(best);
  const uint32_t rr = (uint32_t)(best & 0x1FFFF);
  if (rr == 0x1FFFF) {
    *mx = 0;
    *my = 0;
  } else {
    const uint32_t rank = 0xFFFFu - rr;
    *my = (int)(rank >> 8) - 64;
    *mx = (int)(rank & 255) - 64;
  }
  *minD = (int)(best >> 40);
}

// The scan of table_replay / table_argmin without a predictor, for the
// median-predictor mode's parallel pass.  The predictor enters the integer
// stage only through the tie rule (Full.cpp:424-436): a candidate whose cost
// equals the running minimum is taken unless the incumbent lies strictly
// closer to the rounded predictor.  Here only strictly smaller costs are
// taken, and *tie reports whether the scan ever met an equal cost among the
// candidates the reference would have examined.  Without one, the result is
// the reference's for every predictor; with one, the MB is decided again
// once its predictor is known.
__device__ inline void table_scan_nopred(const TabView& tab, bool filter, int xl, int xr, int yu, int yd,
                                         int lane, int* pmx, int* pmy, int* pminD, int* ptie) {
  const int w = xr - xl + 1, n = w * (yd - yu + 1);
  int mx = 0, my = 0;
  int minD = (int)tab[(0 - yu) * w + (0 - xl)];
  bool tie = false;
  if (!filter) {
    // argmin: the least cost, and whether it is reached more than once
    // (the zero vector's entry included, as the reference's incumbent)
    unsigned long long best = ~0ull;
    int cnt = 0;
    for (int idx = lane; idx < n; idx += 64) {
      const int i = yu + idx / w, j = xl + idx % w;
      const uint32_t c = tab[idx];
      const unsigned long long rr = (i == 0 && j == 0) ? 0x1FFFFull
                                                       : 0xFFFFull - (unsigned long long)((i + 64) * 256 + (j + 64));
      const unsigned long long key = ((unsigned long long)c << 40) | rr;
      best = key < best ? key : best;
    }
    best = wave_min_u64(best);
    const uint32_t cmin = (uint32_t)(best >> 40);
    for (int idx = lane; idx < n; idx += 64) cnt += tab[idx] == cmin;
    tie = wave_sum(cnt) > 1;
    const uint32_t rr = (uint32_t)(best & 0x1FFFF);
    if (rr != 0x1FFFF) {
      const uint32_t rank = 0xFFFFu - rr;
      my = (int)(rank >> 8) - 64;
      mx = (int)(rank & 255) - 64;
    }
    minD = (int)cmin;
  } else {
    int sidx = 0;
    while (sidx < n) {
      const int idx = sidx + lane;
      bool acc = false, eq = false;
      int i = 0, j = 0, c = 0;
      if (idx < n) {
        i = yu + idx / w;
        j = xl + idx % w;
        if ((i | j) != 0 && !((i + my) < yu || (i + my) > yd || (j + mx) < xl || (j + mx) > xr)) {
          c = (int)tab[idx];
          acc = c < minD;
          eq = c == minD;
        }
      }
      const unsigned long long mask = __ballot(acc), eqm = __ballot(eq);
      if (mask == 0) {
        tie |= eqm != 0;
        sidx += 64;
        continue;
      }
      const int f = __ffsll((long long)mask) - 1;
      // equal costs before the accepted one were examined at this minimum
      tie |= (eqm & ((1ull << f) - 1ull)) != 0;
      minD = __builtin_amdgcn_readlane(c, f);
      mx = __builtin_amdgcn_readlane(j, f);
      my = __builtin_amdgcn_readlane(i, f);
      sidx += f + 1;
    }
  }
  *pmx = mx;
  *pmy = my;
  *pminD = minD;
  *ptie = tie ? 1 : 0;
}

// The reference's raster scan with its running-best filter (Full.cpp:401-446).
__device__ inline void table_replay(const TabView& tab, int xl, int xr, int yu, int yd, int pXr, int pYr,
                             int lane, int* pmx, int* pmy, int* pminD) {
  const int w = xr - xl + 1, n = w * (yd - yu + 1);
  int mx = 0, my = 0;
  int minD = (int)tab[(0 - yu) * w + (0 - xl)];
  int sidx = 0;
  while (sidx < n) {
    const int idx = sidx + lane;
    bool acc = false;
    int i = 0, j = 0, c = 0;
    if (idx < n) {
      i = yu + idx / w;
      j = xl + idx % w;
      if ((i | j) != 0 && !((i + my) < yu || (i + my) > yd || (j + mx) < xl || (j + mx) > xr)) {
        c = (int)tab[idx];
        if (c < minD) {
          acc = true;
        } else if (c == minD) {
          const int cx = mx - pXr, cy = my - pYr, nx = j - pXr, ny = i - pYr;
          acc = (cy * cy + cx * cx) - (ny * ny + nx * nx) >= 0;
        }
      }
    }
    const unsigned long long mask = __ballot(acc);
    if (mask == 0) {
      sidx += 64;
      continue;
    }
    // the accepted candidate's lane is wave-uniform: read it out directly
    const int f = __ffsll((long long)mask) - 1;
    minD = __builtin_amdgcn_readlane(c, f);
    mx = __builtin_amdgcn_readlane(j, f);
    my = __builtin_amdgcn_readlane(i, f);
    sidx += f + 1;
  }
  *pmx = mx;
  *pmy = my;
  *pminD = minD;
}

// Build the half-pel lattice around full-pel origin (ox, oy) (patch staged).
// Hg[hy][hx]: even/even G, even/odd b, odd/even h, odd/odd j.  Each lane
// fills whole 2x2 cells (G, b, h, j of one full-pel position), so no lane
// branches on the parity; fixed trip counts let the LDS reads of all
// iterations overlap.
__device__ inline void build_half_grid(WaveLds& s, int lane) {
  constexpr int kHv = 17 * kPatch, kCells = 18 * 18;
#pragma unroll
  for (int j = 0; j < (kHv + 63) / 64; j++) {
    const int i = min(lane + 64 * j, kHv - 1);
    const int a = i / kPatch, x = i - a * kPatch;
    s.hv[i] = (int16_t)vtap_p(s.P, x, a + 2);
  }
  wave_sync();
#pragma unroll
  for (int j = 0; j < (kCells + 63) / 64; j++) {
    const int i = min(lane + 64 * j, kCells - 1);
    const int cy = i / 18, cx = i - cy * 18;
    // (b, h and j of the last row / column are not stored; their reads are
    // clamped into the arrays)
    const int cxi = min(cx, 16), cyi = min(cy, 16);
    const int G = s.P[(cy + 2) * kPatch + cx + 2];
    const int b = b_p(s.P, cxi + 2, cy + 2);
    const int16_t* r = s.hv + cyi * kPatch;
    const int h = clip255((r[cx + 2] + 16) >> 5);
    const int t = ICU_TAP6((int)r[cxi], (int)r[cxi + 1], (int)r[cxi + 2], (int)r[cxi + 3],
                           (int)r[cxi + 4], (int)r[cxi + 5]);
    const int jv = clip255((t + 512) >> 10);
    int16_t* o = s.Hg + (2 * cy) * kHalf + 2 * cx;
    // cells on the last row / column have only their even samples inside
    o[0] = (int16_t)G;
    if (cx < 17) o[1] = (int16_t)b;
    if (cy < 17) {
      o[kHalf] = (int16_t)h;
      if (cx < 17) o[kHalf + 1] = (int16_t)jv;
    }
  }
  wave_sync();
}

// The sub-pel stages without the LDS lattice: lane l owns the 4 pixels of
// block row l / 4, columns 4 (l % 4) .. +3 (its dword cw of the input
// block), whose samples need reference rows y0 + r - 3 .. y0 + r + 3 and
// columns x0 + c0 - 3 .. x0 + c0 + 6.
// This lane's reference pels p[k][i] = (x0 + c0 - 3 + i, y0 + r - 3 + k),
// k = 0..6, i = 0..9 (r = lane / 4, c0 = 4 (lane % 4)), read as 7 x 4
// aligned dwords from the padded plane (L2-resident after the search).
__device__ __forceinline__ void unpack_pels(const uint32_t (&w)[7][4], int sh, int (&p)[7][10]) {
#pragma unroll
  for (int k = 0; k < 7; k++) {
    const uint32_t a0 = __builtin_amdgcn_alignbyte(w[k][1], w[k][0], sh);
    const uint32_t a1 = __builtin_amdgcn_alignbyte(w[k][2], w[k][1], sh);
    const uint32_t a2 = __builtin_amdgcn_alignbyte(w[k][3], w[k][2], sh);
#pragma unroll
    for (int i = 0; i < 10; i++) {
      const uint32_t a = i < 4 ? a0 : (i < 8 ? a1 : a2);
      p[k][i] = (int)((a >> (8 * (i & 3))) & 0xFFu);
    }
  }
}
__device__ __forceinline__ void load_pels(const DecideDesc& d, int x0, int y0, int lane, int (&p)[7][10]) {
  const int r = lane >> 2, c0 = 4 * (lane & 3);
  const int xs = x0 - 3 + c0, sh = xs & 3;  // 0..3 also for negative xs
  const uint32_t* base = (const uint32_t*)(d.ref + (long)(y0 - 3 + r) * d.ref_stride + (xs - sh));
  const int sdw = d.ref_stride >> 2;
  // 10 bytes from byte sh of a row span 4 dwords when sh == 3
  uint32_t w[7][4];
#pragma unroll
  for (int k = 0; k < 7; k++)
#pragma unroll
    for (int q = 0; q < 4; q++) w[k][q] = base[k * sdw + q];
  unpack_pels(w, sh, p);
}
// The eight half-pel candidates' distortions over this lane's 4 pels from
// its reference pels (load_pels): every G/b/h/j sample is filtered in
// registers with the same arithmetic as build_half_grid (vertical 6-tap
// sums first, j from them).  acc[k] receives candidate k (kSubX/kSubY
// order).
template <int METRIC>
__device__ __forceinline__ void half_from_pels(const int (&p)[7][10], uint32_t cw, int (&acc)[8]) {
  // raw vertical sums between rows y0 + r - 1, y0 + r (a = 0) and y0 + r,
  // y0 + r + 1 (a = 1)
  int v[2][10];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int i = 0; i < 10; i++)
      v[a][i] = ICU_TAP6(p[a][i], p[a + 1][i], p[a + 2][i], p[a + 3][i], p[a + 4][i], p[a + 5][i]);
  int B[5], J[2][5], Hh[2][4];
#pragma unroll
  for (int i = 0; i < 5; i++) {
    B[i] = clip255((ICU_TAP6(p[3][i], p[3][i + 1], p[3][i + 2], p[3][i + 3], p[3][i + 4], p[3][i + 5]) + 16) >> 5);
#pragma unroll
    for (int a = 0; a < 2; a++)
      J[a][i] = clip255((ICU_TAP6(v[a][i], v[a][i + 1], v[a][i + 2], v[a][i + 3], v[a][i + 4], v[a][i + 5]) + 512) >> 10);
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int j = 0; j < 4; j++) Hh[a][j] = clip255((v[a][j + 3] + 16) >> 5);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int cv = (int)((cw >> (8 * j)) & 0xFFu);
    acc[0] = dist_add<METRIC>(acc[0], cv, J[0][j]);      // (-1,-1)
    acc[1] = dist_add<METRIC>(acc[1], cv, Hh[0][j]);     // ( 0,-1)
    acc[2] = dist_add<METRIC>(acc[2], cv, J[0][j + 1]);  // ( 1,-1)
    acc[3] = dist_add<METRIC>(acc[3], cv, B[j]);         // (-1, 0)
    acc[4] = dist_add<METRIC>(acc[4], cv, B[j + 1]);     // ( 1, 0)
    acc[5] = dist_add<METRIC>(acc[5], cv, J[1][j]);      // (-1, 1)
    acc[6] = dist_add<METRIC>(acc[6], cv, Hh[1][j]);     // ( 0, 1)
    acc[7] = dist_add<METRIC>(acc[7], cv, J[1][j + 1]);  // ( 1, 1)
  }
}

// The eight half-pel candidates around full-pel (x0, y0) without the LDS
// lattice: lane l owns the 4 pixels of block row l / 4, columns 4 (l % 4)
// .. +3 (its dword cw of the input block).
template <int METRIC>
__device__ __forceinline__ void half_pel_direct(const DecideDesc& d, int x0, int y0, uint32_t cw,
                                                int lane, int (&acc)[8]) {
  int p[7][10];
  load_pels(d, x0, y0, lane, p);
  half_from_pels<METRIC>(p, cw, acc);
}

// floor(v / 2) of a compile-time value
constexpr int floor_half(int v) { return v >= 0 ? v / 2 : -((1 - v) / 2); }

// The half-lattice sample at (Y, X) half pels from this lane's pel jj
// (Y, X in -2..2, known at compile time): G, b, h or j as build_half_grid
// stores them at Hg(2 + 2 r + Y, 2 + 2 c + X), from the lane's reference
// pels and the raw vertical sums V[t + 1] between rows r + t, r + t + 1
// (t = -1, 0).
template <int Y, int X>
__device__ __forceinline__ int lattice_at(const int (&p)[7][10], const int (&V)[2][10], int jj) {
  constexpr int t = floor_half(Y), sx = floor_half(X);
  constexpr bool oy = (Y & 1) != 0, ox = (X & 1) != 0;
  if constexpr (!oy && !ox) {
    return p[3 + t][3 + jj + sx];
  } else if constexpr (!oy) {
    const int* q = p[3 + t] + jj + sx + 1;
    return clip255((ICU_TAP6(q[0], q[1], q[2], q[3], q[4], q[5]) + 16) >> 5);
  } else if constexpr (!ox) {
    return clip255((V[t + 1][3 + jj + sx] + 16) >> 5);
  } else {
    const int* q = V[t + 1] + jj + sx + 1;
    return clip255((ICU_TAP6(q[0], q[1], q[2], q[3], q[4], q[5]) + 512) >> 10);
  }
}

// The eight quarter-pel candidates around the half-pel winner (HX, HY)
// half pels from full pel (x0, y0): the estimator's pairwise averages of
// lattice samples around the winner (LoadQuartPelWindow, as refine_subpel
// reads them from the LDS lattice), here from this lane's reference pels.
template <int HY, int HX>
__device__ __forceinline__ void quarter_from_pels(const int (&p)[7][10], uint32_t cw, int metric, int (&acc)[8]) {
  int V[2][10];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int i = 0; i < 10; i++)
      V[a][i] = ICU_TAP6(p[a][i], p[a + 1][i], p[a + 2][i], p[a + 3][i], p[a + 4][i], p[a + 5][i]);
#pragma unroll
  for (int jj = 0; jj < 4; jj++) {
    const int cv = (int)((cw >> (8 * jj)) & 0xFFu);
    const int uu = lattice_at<HY - 1, HX - 1>(p, V, jj), u0 = lattice_at<HY - 1, HX>(p, V, jj);
    const int l0 = lattice_at<HY, HX - 1>(p, V, jj), c0 = lattice_at<HY, HX>(p, V, jj);
    const int r0 = lattice_at<HY, HX + 1>(p, V, jj), d0 = lattice_at<HY + 1, HX>(p, V, jj);
    acc[0] += pix_dist(cv, (uu + c0 + 1) >> 1, metric);  // (-1,-1)
    acc[1] += pix_dist(cv, (u0 + c0 + 1) >> 1, metric);  // ( 0,-1)
    acc[2] += pix_dist(cv, (u0 + r0 + 1) >> 1, metric);  // ( 1,-1)
    acc[3] += pix_dist(cv, (l0 + c0 + 1) >> 1, metric);  // (-1, 0)
    acc[4] += pix_dist(cv, (r0 + c0 + 1) >> 1, metric);  // ( 1, 0)
    acc[5] += pix_dist(cv, (l0 + d0 + 1) >> 1, metric);  // (-1, 1)
    acc[6] += pix_dist(cv, (d0 + c0 + 1) >> 1, metric);  // ( 0, 1)
    acc[7] += pix_dist(cv, (d0 + r0 + 1) >> 1, metric);  // ( 1, 1)
  }
}

// Reference pels p[k][i] = (x - 3 + i, y - 3 + k) around one pel (x, y)
// of this lane, k = 0..6, i = 0..9, and the raw vertical sums V[a][i]
// between rows y - 1 + a and y + a: what lattice_at needs for jj = 0.
__device__ __forceinline__ void load_pels_at(const DecideDesc& d, int x, int y, int (&p)[7][10], int (&V)[2][10]) {
  const int xs = x - 3, sh = xs & 3;
  const uint32_t* base = (const uint32_t*)(d.ref + (long)(y - 3) * d.ref_stride + (xs - sh));
  const int sdw = d.ref_stride >> 2;
  uint32_t w[7][4];
#pragma unroll
  for (int k = 0; k < 7; k++)
#pragma unroll
    for (int q = 0; q < 4; q++) w[k][q] = base[k * sdw + q];
#pragma unroll
  for (int k = 0; k < 7; k++) {
    const uint32_t a0 = __builtin_amdgcn_alignbyte(w[k][1], w[k][0], sh);
    const uint32_t a1 = __builtin_amdgcn_alignbyte(w[k][2], w[k][1], sh);
    const uint32_t a2 = __builtin_amdgcn_alignbyte(w[k][3], w[k][2], sh);
#pragma unroll
    for (int i = 0; i < 10; i++) {
      const uint32_t a = i < 4 ? a0 : (i < 8 ? a1 : a2);
      p[k][i] = (int)((a >> (8 * (i & 3))) & 0xFFu);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int i = 0; i < 10; i++)
      V[a][i] = ICU_TAP6(p[a][i], p[a + 1][i], p[a + 2][i], p[a + 3][i], p[a + 4][i], p[a + 5][i]);
}

// The six lattice samples the quarter-pel averages take around the half-pel
// winner (HX, HY) for pel jj: (HY - 1, HX - 1), (HY - 1, HX), (HY, HX - 1),
// (HY, HX), (HY, HX + 1), (HY + 1, HX).
template <int HY, int HX>
__device__ __forceinline__ void winner_samples(const int (&p)[7][10], const int (&V)[2][10], int jj, int (&q)[6]) {
  q[0] = lattice_at<HY - 1, HX - 1>(p, V, jj);
  q[1] = lattice_at<HY - 1, HX>(p, V, jj);
  q[2] = lattice_at<HY, HX - 1>(p, V, jj);
  q[3] = lattice_at<HY, HX>(p, V, jj);
  q[4] = lattice_at<HY, HX + 1>(p, V, jj);
  q[5] = lattice_at<HY + 1, HX>(p, V, jj);
}
__device__ __forceinline__ void winner_samples_at(const int (&p)[7][10], const int (&V)[2][10], int jj, int hyh,
                                                  int hxh, int (&q)[6]) {
  switch ((hyh + 1) * 3 + (hxh + 1)) {
    case 0: winner_samples<-1, -1>(p, V, jj, q); break;
    case 1: winner_samples<-1, 0>(p, V, jj, q); break;
    case 2: winner_samples<-1, 1>(p, V, jj, q); break;
    case 3: winner_samples<0, -1>(p, V, jj, q); break;
    case 4: winner_samples<0, 0>(p, V, jj, q); break;
    case 5: winner_samples<0, 1>(p, V, jj, q); break;
    case 6: winner_samples<1, -1>(p, V, jj, q); break;
    case 7: winner_samples<1, 0>(p, V, jj, q); break;
    default: winner_samples<1, 1>(p, V, jj, q); break;
  }
}

// Dispatch on the (wave-uniform) half-pel winner.
__device__ __forceinline__ void quarter_direct(const int (&p)[7][10], uint32_t cw, int metric, int hyh, int hxh,
                                               int (&acc)[8]) {
  switch ((hyh + 1) * 3 + (hxh + 1)) {
    case 0: quarter_from_pels<-1, -1>(p, cw, metric, acc); break;
    case 1: quarter_from_pels<-1, 0>(p, cw, metric, acc); break;
    case 2: quarter_from_pels<-1, 1>(p, cw, metric, acc); break;
    case 3: quarter_from_pels<0, -1>(p, cw, metric, acc); break;
    case 4: quarter_from_pels<0, 0>(p, cw, metric, acc); break;
    case 5: quarter_from_pels<0, 1>(p, cw, metric, acc); break;
    case 6: quarter_from_pels<1, -1>(p, cw, metric, acc); break;
    case 7: quarter_from_pels<1, 0>(p, cw, metric, acc); break;
    default: quarter_from_pels<1, 1>(p, cw, metric, acc); break;
  }
}

// The SAD form of half_pel_direct in packed 16-bit arithmetic: the same
// samples and the same distortions, with two columns in one register.
// Pairs (p[k][i], p[k][i + 1]) are cut straight out of the loaded dwords by
// v_perm (bytes sh + i and sh + i + 1 of dwords i / 4 and i / 4 + 1, each
// zero-extended to 16 bits); the vertical sums (-2550 .. 10710) and the
// G/b/h samples stay in 16-bit halves, the j samples (up to 450k before
// rounding) come from three v_dot2_i32_i16 each, and every distortion
// is a v_sad_u16 against a pair of block pels.
typedef short icu_s16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ icu_s16x2 pk_s16(uint32_t x) { return __builtin_bit_cast(icu_s16x2, x); }
__device__ __forceinline__ uint32_t pk_u32(icu_s16x2 x) { return __builtin_bit_cast(uint32_t, x); }

__device__ __forceinline__ icu_s16x2 pk_tap6(icu_s16x2 a, icu_s16x2 b, icu_s16x2 c, icu_s16x2 d, icu_s16x2 e,
                                             icu_s16x2 f) {
  return (a + f) + (c + d) * (icu_s16x2){20, 20} + (b + e) * (icu_s16x2){-5, -5};
}

// clip255((x + 16) >> 5) in each half
__device__ __forceinline__ icu_s16x2 pk_round5(icu_s16x2 x) {
  const icu_s16x2 r = (x + (icu_s16x2){16, 16}) >> (icu_s16x2){5, 5};
  return __builtin_elementwise_min(__builtin_elementwise_max(r, (icu_s16x2){0, 0}), (icu_s16x2){255, 255});
}

// clip255((TAP6(v[i .. i + 5]) + 512) >> 10) from the pairs (v[i], v[i + 1]),
// (v[i + 2], v[i + 3]), (v[i + 4], v[i + 5])
__device__ __forceinline__ uint32_t pk_j(icu_s16x2 q0, icu_s16x2 q2, icu_s16x2 q4) {
  int t = __builtin_amdgcn_sdot2(q4, (icu_s16x2){-5, 1}, 512, false);
  t = __builtin_amdgcn_sdot2(q2, (icu_s16x2){20, 20}, t, false);
  t = __builtin_amdgcn_sdot2(q0, (icu_s16x2){1, -5}, t, false);
  return (uint32_t)clip255(t >> 10);
}

__device__ __forceinline__ void half_pel_sad_pk(const DecideDesc& d, int x0, int y0, uint32_t cw, int lane,
                                                int (&acc)[8]) {
  const int r = lane >> 2, c0 = 4 * (lane & 3);
  // the byte offset of column x0 - 3 + c0 in its dword, the same in every
  // lane (c0 is a multiple of 4; 0..3 also for negative columns); rows are
  // read at 32-bit offsets from the wave's origin
  const int sh = (x0 - 3) & 3;
  typedef __attribute__((address_space(1))) const uint8_t g_u8;
  typedef __attribute__((address_space(1))) const uint32_t g_u32;
  g_u8* org = (g_u8*)d.ref + (long)(y0 - 3) * d.ref_stride + (x0 - 3 - sh);
  const uint32_t off = (uint32_t)(r * d.ref_stride + c0);
  uint32_t w[7][4];
#pragma unroll
  for (int k = 0; k < 7; k++) {
    g_u32* row = (g_u32*)(org + (off + (uint32_t)(k * d.ref_stride)));
#pragma unroll
    for (int q = 0; q < 4; q++) w[k][q] = row[q];
  }
  // selector of the pair at i with i % 4 == t: bytes sh + t, sh + t + 1 of
  // (w[i / 4], w[i / 4 + 1]), each followed by a zero byte
  uint32_t sel[4];
#pragma unroll
  for (int t = 0; t < 4; t++) sel[t] = 0x0c010c00u + (uint32_t)(sh + t) * 0x00010001u;
  // row pairs (p[k][2e], p[k][2e + 1]) and their vertical sums
  icu_s16x2 P[7][5];
#pragma unroll
  for (int k = 0; k < 7; k++)
#pragma unroll
    for (int e = 0; e < 5; e++) P[k][e] = pk_s16(__builtin_amdgcn_perm(w[k][e / 2 + 1], w[k][e / 2], sel[(2 * e) & 3]));
  // VP[a][i] = (v[a][i], v[a][i + 1]), i = 0..8
  icu_s16x2 VP[2][9];
#pragma unroll
  for (int a = 0; a < 2; a++) {
#pragma unroll
    for (int e = 0; e < 5; e++)
      VP[a][2 * e] = pk_tap6(P[a][e], P[a + 1][e], P[a + 2][e], P[a + 3][e], P[a + 4][e], P[a + 5][e]);
#pragma unroll
    for (int e = 0; e < 4; e++)
      VP[a][2 * e + 1] = pk_s16(__builtin_amdgcn_alignbyte(pk_u32(VP[a][2 * e + 2]), pk_u32(VP[a][2 * e]), 2));
  }
  // row y0 + r: every pair (p[3][i], p[3][i + 1]), i = 0..9, and the b
  // samples in pairs (B[0], B[1]), (B[2], B[3]), (B[4], -)
  icu_s16x2 R3[10];
#pragma unroll
  for (int i = 0; i < 10; i++) R3[i] = pk_s16(__builtin_amdgcn_perm(w[3][i / 4 + 1], w[3][i / 4], sel[i & 3]));
  icu_s16x2 PB[3];
#pragma unroll
  for (int e = 0; e < 3; e++)
    PB[e] = pk_round5(pk_tap6(R3[2 * e], R3[2 * e + 1], R3[2 * e + 2], R3[2 * e + 3], R3[2 * e + 4], R3[2 * e + 5]));
  const uint32_t b12 = __builtin_amdgcn_alignbyte(pk_u32(PB[1]), pk_u32(PB[0]), 2);
  const uint32_t b34 = __builtin_amdgcn_alignbyte(pk_u32(PB[2]), pk_u32(PB[1]), 2);
  // the block pels (c0, c1), (c2, c3) of this lane
  const uint32_t C0 = __builtin_amdgcn_perm(0u, cw, 0x0c010c00u), C1 = __builtin_amdgcn_perm(0u, cw, 0x0c030c02u);
#pragma unroll
  for (int a = 0; a < 2; a++) {
    uint32_t J[5];
#pragma unroll
    for (int i = 0; i < 5; i++) J[i] = pk_j(VP[a][i], VP[a][i + 2], VP[a][i + 4]);
    const uint32_t h01 = pk_u32(pk_round5(VP[a][3])), h23 = pk_u32(pk_round5(VP[a][5]));
    int& aj0 = acc[a ? 5 : 0];  // (-1, -1 + 2a)
    int& ah = acc[a ? 6 : 1];   // ( 0, -1 + 2a)
    int& aj1 = acc[a ? 7 : 2];  // ( 1, -1 + 2a)
    aj0 = (int)__builtin_amdgcn_sad_u16(J[0] | (J[1] << 16), C0, (uint32_t)aj0);
    aj0 = (int)__builtin_amdgcn_sad_u16(J[2] | (J[3] << 16), C1, (uint32_t)aj0);
    ah = (int)__builtin_amdgcn_sad_u16(h01, C0, (uint32_t)ah);
    ah = (int)__builtin_amdgcn_sad_u16(h23, C1, (uint32_t)ah);
    aj1 = (int)__builtin_amdgcn_sad_u16(J[1] | (J[2] << 16), C0, (uint32_t)aj1);
    aj1 = (int)__builtin_amdgcn_sad_u16(J[3] | (J[4] << 16), C1, (uint32_t)aj1);
  }
  acc[3] = (int)__builtin_amdgcn_sad_u16(pk_u32(PB[0]), C0, (uint32_t)acc[3]);  // (-1, 0)
  acc[3] = (int)__builtin_amdgcn_sad_u16(pk_u32(PB[1]), C1, (uint32_t)acc[3]);
  acc[4] = (int)__builtin_amdgcn_sad_u16(b12, C0, (uint32_t)acc[4]);  // (1, 0)
  acc[4] = (int)__builtin_amdgcn_sad_u16(b34, C1, (uint32_t)acc[4]);
}

// The half- and quarter-pel stages of refine_subpel (subpel >= 2) from
// this lane's reference pels around the full-pel winner (load_pels): the
// eight half-pel candidates (strictly better only),
// then the eight quarter-pel averages around the half-pel winner.  mvx /
// mvy hold the full-pel vector in quarter pels on entry.
__device__ __forceinline__ void subpel_from_pels(const DecideDesc& d, const int (&p)[7][10], uint32_t cw, int& minD,
                                                 int& mvx, int& mvy) {
  int hxh = 0, hyh = 0;
  {
    int acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (d.metric == kMetricSAD) {
      half_from_pels<kMetricSAD>(p, cw, acc);
      wave_sum8_small(acc);
    } else {
      half_from_pels<kMetricSSD>(p, cw, acc);
      wave_sum8(acc);
    }
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (acc[k] < minD) {
        minD = acc[k];
        hxh = kSubX[k];
        hyh = kSubY[k];
      }
  }
  int qmx = 2 * hxh, qmy = 2 * hyh;
  {
    const int hx = 2 * hxh, hy = 2 * hyh;
    int acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    quarter_direct(p, cw, d.metric, hyh, hxh, acc);
    if (d.metric == kMetricSAD) wave_sum8_small(acc);
    else wave_sum8(acc);
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (acc[k] < minD) {
        minD = acc[k];
        qmx = hx + kSubX[k];
        qmy = hy + kSubY[k];
      }
  }
  mvx += qmx;
  mvy += qmy;
}

// Half-pel then (subpel >= 2) quarter-pel refinement around full-pel
// (mx, my) of the MB at (n, m), strictly better distortions only
// (Full.cpp:452-517; UMHS.cpp:795-862 is the same stage).  cw: this lane's
// dword of the input block (row lane / 4, columns 4 (lane % 4)), also
// staged in s.cur4.  Updates minD; returns the vector in quarter pels.
// Quarter pel: LATTICE builds the half-pel lattice in LDS (s.P, s.hv, s.Hg;
// for kernels whose VGPRs are spoken for), else every sample is filtered in
// registers from the padded plane (no LDS, fewer dependent steps).
template <bool LATTICE = false>
__device__ inline void refine_subpel(const DecideDesc& d, WaveLds& s, int n, int m, int mx, int my, uint32_t cw,
                                     int lane, int& minD, int& mvx, int& mvy) {
  mvx = mx * 4;
  mvy = my * 4;
  if (d.subpel == 1) {
    // half pel only: the eight candidates filtered in registers
    int acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (d.metric == kMetricSAD) {
      half_pel_sad_pk(d, n + mx, m + my, cw, lane, acc);
      wave_sum8_small(acc);
    } else {
      half_pel_direct<kMetricSSD>(d, n + mx, m + my, cw, lane, acc);
      wave_sum8(acc);
    }
    int hxh = 0, hyh = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (acc[k] < minD) {
        minD = acc[k];
        hxh = kSubX[k];
        hyh = kSubY[k];
      }
    mvx += 2 * hxh;
    mvy += 2 * hyh;
  } else if (d.subpel >= 2 && LATTICE) {
    // the 35 x 35 half-pel lattice built in LDS from the reference patch
    const uint8_t* cur8 = (const uint8_t*)s.cur4;
    load_patch(d, n + mx, m + my, s.P, lane);
    wave_sync();
    build_half_grid(s, lane);
    // half-pel candidates, all eight summed at once, then taken in the
    // reference's order (strictly better only)
    int hxh = 0, hyh = 0;
    {
      int acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int p = lane; p < 256; p += 64) {
        const int r = p >> 4, c = p & 15;
        const int cv = cur8[p];
        const int16_t* h = s.Hg + (2 + 2 * r) * kHalf + 2 + 2 * c;
#pragma unroll
        for (int k = 0; k < 8; k++) acc[k] += pix_dist(cv, h[kSubY[k] * kHalf + kSubX[k]], d.metric);
      }
      wave_sum8(acc);
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (acc[k] < minD) {
          minD = acc[k];
          hxh = kSubX[k];
          hyh = kSubY[k];
        }
    }
    int qmx = 2 * hxh, qmy = 2 * hyh;
    if (d.subpel >= 2) {
      // quarter candidates: the reference's pairwise averages of half-grid
      // samples around the half-pel winner H = (cr, cc) (LoadQuartPelWindow)
      const int hx = 2 * hxh, hy = 2 * hyh;
      int acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int p = lane; p < 256; p += 64) {
        const int r = p >> 4, c = p & 15;
        const int cv = cur8[p];
        const int16_t* h = s.Hg + (2 + 2 * r + hyh) * kHalf + 2 + 2 * c + hxh;
        const int uu = h[-kHalf - 1], u0 = h[-kHalf], l0 = h[-1], c0 = h[0], r0 = h[1], d0 = h[kHalf];
        acc[0] += pix_dist(cv, (uu + c0 + 1) >> 1, d.metric);  // (-1,-1)
        acc[1] += pix_dist(cv, (u0 + c0 + 1) >> 1, d.metric);  // ( 0,-1)
        acc[2] += pix_dist(cv, (u0 + r0 + 1) >> 1, d.metric);  // ( 1,-1)
        acc[3] += pix_dist(cv, (l0 + c0 + 1) >> 1, d.metric);  // (-1, 0)
        acc[4] += pix_dist(cv, (r0 + c0 + 1) >> 1, d.metric);  // ( 1, 0)
        acc[5] += pix_dist(cv, (l0 + d0 + 1) >> 1, d.metric);  // (-1, 1)
        acc[6] += pix_dist(cv, (d0 + c0 + 1) >> 1, d.metric);  // ( 0, 1)
        acc[7] += pix_dist(cv, (d0 + r0 + 1) >> 1, d.metric);  // ( 1, 1)
      }
      wave_sum8(acc);
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (acc[k] < minD) {
          minD = acc[k];
          qmx = hx + kSubX[k];
          qmy = hy + kSubY[k];
        }
    }
    mvx += qmx;
    mvy += qmy;
    wave_sync();
  } else if (d.subpel >= 2) {
    // both stages from this lane's reference pels, filtered in registers
    int p[7][10];
    load_pels(d, n + mx, m + my, lane, p);
    subpel_from_pels(d, p, cw, minD, mvx, mvy);
    wave_sync();
  }

}

// Four quarter-pel samples (OverlayMem2Dv2::QuarterRead's 16 cases, H.264
// 8.4.2.2.1) at full pels (x .. x + 3, y) with fractions fx, fy in 0..3,
// filtered in registers from reference rows y - 2 .. y + 3 and columns
// x - 2 .. x + 6 of the padded plane, read as 6 x 3 aligned dwords.
__device__ __forceinline__ void qpel_row4(const uint8_t* ref, int stride, int x, int y, int fx, int fy,
                                          int (&out)[4]) {
  const int xs = x - 2, xa = xs & ~3, sh = xs - xa;
  const uint32_t* base = (const uint32_t*)(ref + (long)(y - 2) * stride + xa);
  const int sdw = stride >> 2;
  uint32_t d[6][3];
#pragma unroll
  for (int k = 0; k < 6; k++)
#pragma unroll
    for (int q = 0; q < 3; q++) d[k][q] = base[k * sdw + q];
  // p[k][c]: reference (x - 2 + c, y - 2 + k); the full pels are p[2][2 + c]
  int p[6][9];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const uint32_t a0 = __builtin_amdgcn_alignbyte(d[k][1], d[k][0], sh);
    const uint32_t a1 = __builtin_amdgcn_alignbyte(d[k][2], d[k][1], sh);
    const uint32_t a2 = d[k][2] >> (8 * sh);
#pragma unroll
    for (int c = 0; c < 9; c++) {
      const uint32_t a = c < 4 ? a0 : (c < 8 ? a1 : a2);
      p[k][c] = (int)((a >> (8 * (c & 3))) & 0xFFu);
    }
  }
  auto G = [&](int r, int c) { return p[2 + r][2 + c]; };
  auto B = [&](int r, int c) {  // right of (c, r), r = 0 or 1
    const int* q = p[2 + r];
    return clip255((ICU_TAP6(q[c], q[c + 1], q[c + 2], q[c + 3], q[c + 4], q[c + 5]) + 16) >> 5);
  };
  auto V = [&](int cc) {  // raw vertical sum below row 0 at column cc - 2
    return ICU_TAP6(p[0][cc], p[1][cc], p[2][cc], p[3][cc], p[4][cc], p[5][cc]);
  };
  auto Hh = [&](int c) { return clip255((V(2 + c) + 16) >> 5); };
  auto J = [&](int c) {
    return clip255((ICU_TAP6(V(c), V(c + 1), V(c + 2), V(c + 3), V(c + 4), V(c + 5)) + 512) >> 10);
  };
#define ICU_ROW4(expr) _Pragma("unroll") for (int c = 0; c < 4; c++) out[c] = (expr)
  switch (fx | (fy << 2)) {
    case 0: ICU_ROW4(G(0, c)); break;
    case 2: ICU_ROW4(B(0, c)); break;
    case 8: ICU_ROW4(Hh(c)); break;
    case 10: ICU_ROW4(J(c)); break;
    case 1: ICU_ROW4((G(0, c) + B(0, c) + 1) >> 1); break;
    case 3: ICU_ROW4((G(0, c + 1) + B(0, c) + 1) >> 1); break;
    case 4: ICU_ROW4((G(0, c) + Hh(c) + 1) >> 1); break;
    case 12: ICU_ROW4((G(1, c) + Hh(c) + 1) >> 1); break;
    case 6: ICU_ROW4((J(c) + B(0, c) + 1) >> 1); break;
    case 14: ICU_ROW4((J(c) + B(1, c) + 1) >> 1); break;
    case 9: ICU_ROW4((J(c) + Hh(c) + 1) >> 1); break;
    case 11: ICU_ROW4((J(c) + Hh(c + 1) + 1) >> 1); break;
    case 5: ICU_ROW4((B(0, c) + Hh(c) + 1) >> 1); break;
    case 7: ICU_ROW4((B(0, c) + Hh(c + 1) + 1) >> 1); break;
    case 13: ICU_ROW4((Hh(c) + B(1, c) + 1) >> 1); break;
    default: ICU_ROW4((Hh(c + 1) + B(1, c) + 1) >> 1); break;
  }
#undef ICU_ROW4
}

// Distortion of the input block against the quarter-pel read at full pel
// (x0, y0) + fractions (fx, fy) in -3..3 (OverlayMem2Dv2::QuarterRead; a
// negative fraction reads from the pel to the left / above with 4 + f,
// OverlayMem2Dv2.cpp:709-724).
__device__ inline int quarter_dist(const DecideDesc& d, WaveLds& s, int x0, int y0, int fx, int fy, int lane) {
  int fox = 0, foy = 0;
  if (fx < 0) { fox = -1; fx = 4 + fx; }
  if (fy < 0) { foy = -1; fy = 4 + fy; }
  // lane l: block row l / 4, columns 4 (l % 4) .. + 3
  const int r = lane >> 2, c0 = 4 * (lane & 3);
  int q[4];
  qpel_row4(d.ref, d.ref_stride, x0 + fox + c0, y0 + foy + r, fx, fy, q);
  const uint32_t cw = s.cur4[lane];
  int acc = 0;
#pragma unroll
  for (int c = 0; c < 4; c++) acc += pix_dist((int)((cw >> (8 * c)) & 0xFF), q[c], d.metric);
  return wave_sum(acc);
}

struct Decision {
  int mvx, mvy, dist;
  int imx, imy, iminD;
};

__device__ inline Decision finish_mb(const DecideDesc& d, WaveLds& s, int mb, int predX, int predY, int lane,
                                     int mvx, int mvy, int minD, int zeroD, Decision out);

// The median-predictor mode's predictor-free pass over one MB: the integer
// stage without a predictor (table_scan_nopred) and the sub-pel refinement,
// which depends on the integer vector only.
struct PredFree {
  int mvx, mvy, dist;   // refined vector (quarter pel) and its distortion
  int imx, imy, iminD;  // integer stage
  int zeroD;            // the zero vector's cost
  int tie;              // the scan met an equal cost: decide again with the predictor (last field:
                        // icu_me_tie_count reads it there)
};

// decide_mb for the lean modes: SAD, no cost table, integer or half-pel
// refinement, no final vector choice (the zero predictor's GetBestVector is
// off).  The same outputs as decide_mb there, from far less code: the wave
// kernels that take only such streams carry this instead, which keeps them
// inside the instruction cache.
__device__ __forceinline__ void decide_mb_lean(const DecideDesc& d, int mb, int lane, const IntResult& r) {
  const int col = mb % d.mbw, row = mb / d.mbw;
  const int n = col * 16, m = row * 16;
  const int slot = d.mb_base + mb;
  int minD = r.cost, mvx = r.mx * 4, mvy = r.my * 4;
  if (d.subpel == 1) {
    typedef __attribute__((address_space(1))) const uint8_t g_u8;
    const uint32_t cw = *(__attribute__((address_space(1))) const uint32_t*)(
        (g_u8*)d.cur + (long)m * d.cur_stride + n + (uint32_t)((lane >> 2) * d.cur_stride + 4 * (lane & 3)));
    int acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    half_pel_sad_pk(d, n + r.mx, m + r.my, cw, lane, acc);
    wave_sum8_small(acc);
    int hxh = 0, hyh = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (acc[k] < minD) {
        minD = acc[k];
        hxh = kSubX[k];
        hyh = kSubY[k];
      }
    mvx += 2 * hxh;
    mvy += 2 * hyh;
  }
  if (lane == 0) {
    const uint32_t w = (uint32_t)(uint16_t)(int16_t)mvx | ((uint32_t)(uint16_t)(int16_t)mvy << 16);
    __hip_atomic_store((uint32_t*)(d.mv_out + 2 * slot), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    d.dist_out[slot] = minD;
    if (d.int_out) {
      d.int_out[3 * slot] = r.mx;
      d.int_out[3 * slot + 1] = r.my;
      d.int_out[3 * slot + 2] = r.cost;
    }
  }
}

// staged_tab (nullable): the MB's cost table already copied to LDS.
template <bool LATTICE = false>
__device__ inline Decision decide_mb(const DecideDesc& d, WaveLds& s, int mb, int predX, int predY,
                                     int lane, const IntResult* known = nullptr,
                                     const uint32_t* staged_tab = nullptr) {
  const int col = mb % d.mbw, row = mb / d.mbw;
  const int n = col * 16, m = row * 16;
  const int slot = d.mb_base + mb;
  // input block: one aligned dword per lane (current planes are 16-byte
  // aligned with a stride that is a multiple of 16)
  const uint32_t cw = *(const uint32_t*)(d.cur + (long)(m + (lane >> 2)) * d.cur_stride + n + 4 * (lane & 3));
  s.cur4[lane] = cw;
  // the zero vector's cost, which the argmin search does not report; only
  // GetBestVector reads it (the padded plane's rows are 16-byte aligned)
  int zero_direct = 0;
  if (!d.table && d.bestvec) {
    const uint32_t rw = *(const uint32_t*)(d.ref + (long)(m + (lane >> 2)) * d.ref_stride + n + 4 * (lane & 3));
    int z = 0;
    if (d.metric == kMetricSAD) {
      z = (int)__builtin_amdgcn_sad_u8(cw, rw, 0u);  // the 4 byte differences at once
    } else {
#pragma unroll
      for (int i = 0; i < 4; i++) z += pix_dist((int)((cw >> (8 * i)) & 0xFF), (int)((rw >> (8 * i)) & 0xFF), d.metric);
    }
    zero_direct = wave_sum(z);
  }

  const int predXQ = predX % 4, predYQ = predY % 4;
  int predX0 = predX / 4, predY0 = predY / 4;
  const int pXr = predX < 0 ? (predX - 2) / 4 : (predX + 2) / 4;
  const int pYr = predY < 0 ? (predY - 2) / 4 : (predY + 2) / 4;

  int mx, my, minD, zeroD;
  int xl, xr, yu, yd;
  motion_window(n, m, d.W, d.H, d.L, &xl, &xr, &yu, &yd);
  if (d.table) {
    TabView tab;
    tab.t = staged_tab ? staged_tab : d.table + d.table_base + (long long)mb * d.table_dim * d.table_dim;
    tab.dim = d.table_dim;
    tab.L = d.L;
    tab.xl = xl;
    tab.yu = yu;
    tab.w = xr - xl + 1;
    zeroD = (int)tab[(0 - yu) * tab.w + (0 - xl)];
    if (d.filter) table_replay(tab, xl, xr, yu, yd, pXr, pYr, lane, &mx, &my, &minD);
    else table_argmin(tab, xl, xr, yu, yd, pXr, pYr, lane, &mx, &my, &minD);
  } else {
    const IntResult r = known ? *known : d.ires[slot];
    mx = r.mx;
    my = r.my;
    minD = r.cost;
    zeroD = zero_direct;
  }
  Decision out;
  out.imx = mx;
  out.imy = my;
  out.iminD = minD;

  int mvx, mvy;
  refine_subpel<LATTICE>(d, s, n, m, mx, my, cw, lane, minD, mvx, mvy);
  return finish_mb(d, s, mb, predX, predY, lane, mvx, mvy, minD, zeroD, out);
}

// The predictor-dependent end of an MB's decision: the predictor's
// distortion and GetBestVector over {refined best, predictor, zero}
// (Full.cpp:529-579), then the outputs.  out carries the integer stage.
__device__ inline Decision finish_mb(const DecideDesc& d, WaveLds& s, int mb, int predX, int predY, int lane,
                                     int mvx, int mvy, int minD, int zeroD, Decision out) {
  const int col = mb % d.mbw, row = mb / d.mbw;
  const int n = col * 16, m = row * 16;
  const int slot = d.mb_base + mb;
  const int predXQ = predX % 4, predYQ = predY % 4;
  int predX0 = predX / 4, predY0 = predY / 4;
  int outx = mvx, outy = mvy, outd = minD;
  if (d.bestvec) {
    if (predX0 + n > d.W) predX0 = 16;
    if (predX0 + n < -16) predX0 = -16;
    if (predY0 + m > d.H) predY0 = 16;
    if (predY0 + m < -16) predY0 = -16;
    const int pX = predX0 * 4 + predXQ, pY = predY0 * 4 + predYQ;
    int predD;
    if (predXQ == 0 && predYQ == 0 && predX0 == 0 && predY0 == 0) {
      predD = zeroD;
    } else {
      predD = quarter_dist(d, s, predX0 + n, predY0 + m, predXQ, predYQ, lane);
    }
    // fifo [best, pred, zero, zero, zero]; strictly smaller cost wins
    // (selects rather than an indexed array, which would live in scratch)
    double cc = best_vector_cost(minD, (mvx - pX) * (mvx - pX) + (mvy - pY) * (mvy - pY));
    const double cp = best_vector_cost(predD, 0);
    if (cp < cc) {
      cc = cp;
      outx = pX;
      outy = pY;
      outd = predD;
    }
    const double cz = best_vector_cost(zeroD, pX * pX + pY * pY);
    if (cz < cc) {
      outx = 0;
      outy = 0;
      outd = zeroD;
    }
  }
  out.mvx = outx;
  out.mvy = outy;
  out.dist = outd;
  if (lane == 0) {
    const uint32_t w = (uint32_t)(uint16_t)(int16_t)outx | ((uint32_t)(uint16_t)(int16_t)outy << 16);
    __hip_atomic_store((uint32_t*)(d.mv_out + 2 * slot), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    d.dist_out[slot] = outd;
    if (d.int_out) {
      d.int_out[3 * slot] = out.imx;
      d.int_out[3 * slot + 1] = out.imy;
      d.int_out[3 * slot + 2] = out.iminD;
    }
  }
  return out;
}

// The predictor-free pass of one MB (cost-table modes).
__device__ inline void predfree_mb(const DecideDesc& d, WaveLds& s, int mb, int lane, PredFree* out) {
  const int col = mb % d.mbw, row = mb / d.mbw;
  const int n = col * 16, m = row * 16;
  const uint32_t cw = *(const uint32_t*)(d.cur + (long)(m + (lane >> 2)) * d.cur_stride + n + 4 * (lane & 3));
  s.cur4[lane] = cw;
  int xl, xr, yu, yd;
  motion_window(n, m, d.W, d.H, d.L, &xl, &xr, &yu, &yd);
  TabView tab;
  tab.t = d.table + d.table_base + (long long)mb * d.table_dim * d.table_dim;
  tab.dim = d.table_dim;
  tab.L = d.L;
  tab.xl = xl;
  tab.yu = yu;
  tab.w = xr - xl + 1;
  const int zeroD = (int)tab[(0 - yu) * tab.w + (0 - xl)];
  int mx, my, minD, tie;
  table_scan_nopred(tab, d.filter != 0, xl, xr, yu, yd, lane, &mx, &my, &minD, &tie);
  PredFree pf;
  pf.imx = mx;
  pf.imy = my;
  pf.iminD = minD;
  pf.zeroD = zeroD;
  pf.tie = tie;
  int mvx, mvy;
  refine_subpel(d, s, n, m, mx, my, cw, lane, minD, mvx, mvy);
  pf.mvx = mvx;
  pf.mvy = mvy;
  pf.dist = minD;
  if (lane == 0) *out = pf;
}

// An MB's decision once its predictor is known, from its predictor-free
// pass: without a tie only the predictor-dependent end remains.
// cw: this lane's dword of the input block (row lane / 4, columns 4 (lane % 4)).
__device__ inline Decision decide_mb_after(const DecideDesc& d, WaveLds& s, int mb, int predX, int predY, int lane,
                                           const PredFree& pf, uint32_t cw, const uint32_t* staged_tab = nullptr) {
  if (pf.tie) return decide_mb(d, s, mb, predX, predY, lane, nullptr, staged_tab);
  s.cur4[lane] = cw;
  Decision out;
  out.imx = pf.imx;
  out.imy = pf.imy;
  out.iminD = pf.iminD;
  return finish_mb(d, s, mb, predX, predY, lane, pf.mvx, pf.mvy, pf.dist, pf.zeroD, out);
}

// Hand-off between the waves of a wavefront kernel.  Everything one wave
// reads of another's results in the same launch (vectors, distortions, the
// FHS fill records, the progress counters) goes through relaxed agent-scope
// atomics, which the hardware keeps coherent across the XCDs by themselves.
// So a progress counter (FHS) needs only the producer's own stores drained
// (s_waitcnt) before its store, and the consumer only a relaxed poll; a
// hand-off word that carries the result itself (the decision, UMHS and
// Cross wavefronts) needs no drain at all: the consumer polls that word
// until it is no longer kVecPending.  An
// acquire / release at agent scope would add an L2 invalidate to every poll
// and an L2 write-back to every MB, for data that no plain load reads.
// A vector word not yet decided, the median wavefronts' hand-off value:
// (-32768, -32768) quarter pel, which no MB can get (every vector stays
// within 4 (range + 1) of zero).
constexpr uint32_t kVecPending = 0x80008000u;
__device__ __forceinline__ int poll_progress(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void publish_progress(int* p, int v) {
  __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have landed
  asm volatile("" ::: "memory");
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int median3(int x, int y, int z) {
  int mn, mx;
  if (y - z < 0) { mn = y; mx = z; } else { mn = z; mx = y; }
  if (x < mn) mn = x;
  if (x > mx) mx = x;
  return x + y + z - mn - mx;
}

}  // namespace dec
}  // namespace icu
