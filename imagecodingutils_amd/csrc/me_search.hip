// Integer full-search kernels for 16x16 macroblocks on gfx950.
//
// Restates the integer stage of MotionEstimatorH264ImplFull::Estimate
// (src/CodecUtils/MotionEstimatorH264ImplFull.cpp:383-446) with the costs of
// OverlayMem2Dv2::Tad16x16 / Tsd16x16 (src/ImageUtils/OverlayMem2Dv2.cpp:
// 1623-1695, 2655-2735) and the window of GetMotionRange (:1029-1081).
//
// Work decomposition.  A workgroup task is G consecutive MBs of one MB row.
// The union of their search windows is staged once in LDS from the padded
// 8-bit reference (dword rows, read coalesced), their 16x16 blocks beside
// it.  Each lane then owns a tile of candidates: 4 adjacent columns (one
// dword-aligned quad) x V adjacent rows.  It keeps its MB's block in 64
// VGPRs, walks the V+15 reference rows its tile touches once each, and for
// every (reference row, block row) pair issues 4 v_qsad_pk_u16_u8: each
// compares one block dword with an 8-byte reference span and accumulates the
// SADs of the 4 column offsets at once into packed 16-bit sums (a 16x16 SAD
// is at most 65280, so u16 never wraps).  SSD uses v_dot4_u32_u8:
//   sum (c - r)^2 = sum r^2 - 2 sum c*r  (+ sum c^2, added once per MB).
// The per-MB decision is a 64-bit atomic min in LDS over packed keys that
// order candidates exactly as the reference's scan does with its running-
// best filter switched off; optionally every cost is written to a table for
// the decision kernels that replay the filtered scan or use a predictor.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "icu_device.h"
#include "icu_internal.h"
#include "icu_decide.h"

namespace icu {

constexpr int kThreads = 256;
// One staging buffer (window + blocks); the kernel holds two, and the pair
// plus the static arrays stays under the 64 KB dynamic-LDS launch limit.
constexpr int kWinDwords = 7680;
constexpr int kMaxG = 16;         // MBs per task

// LDS-qualified views of the staging buffers: reads through them compile to
// ds_read (a generic pointer picked at run time would become a flat load).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;

// ---------------------------------------------------------------- padding
// Edge-replicated copy of an 8-bit plane into the padded plane
// (OverlayExtMem2Dv2::ExtendBoundary/FillBoundary,
// src/ImageUtils/OverlayExtMem2Dv2.cpp:120-217, as a coordinate clamp).
// 16 bytes per thread.  kPad is a multiple of 16, so a chunk lies wholly
// inside the picture's columns (one 16-byte load of the clamped row), wholly
// left / right of it (the row's first / last pel, one byte load,
// replicated), or, only where W is not a multiple of 16 (4:2:0 chroma
// planes), across the right edge (bytes gathered with the column clamped).
// Every plane of a batch pads in one launch: a workgroup finds its plane by
// binary search over the planes' first workgroups.
__device__ __forceinline__ void pad_chunk(const PadDesc& p) {
  const int cpr = (p.W + 2 * kPad) / 16;  // 16-byte chunks per padded row
  const int idx = ((int)blockIdx.x - p.wg_base) * (int)blockDim.x + (int)threadIdx.x;
  if (idx >= cpr * (p.H + 2 * kPad)) return;
  const int y = idx / cpr, x16 = (idx - y * cpr) * 16;
  const int yy = min(max(y - kPad, 0), p.H - 1);
  const uint8_t* row = p.src + (long)yy * p.src_stride;
  const int x = x16 - kPad;
  uint4 v;
  if (x >= 0 && x + 16 <= p.W) {
    if (((uintptr_t)row & 15) == 0) {
      v = *(const uint4*)(row + x);
    } else {
      // a caller's plane without 16-byte aligned rows
      uint32_t w[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (((uintptr_t)row & 3) == 0) {
          w[q] = *(const uint32_t*)(row + x + 4 * q);
        } else {
          w[q] = 0;
#pragma unroll
          for (int i = 0; i < 4; i++) w[q] |= (uint32_t)row[x + 4 * q + i] << (8 * i);
        }
      }
      v = make_uint4(w[0], w[1], w[2], w[3]);
    }
  } else if (x < 0 || x >= p.W) {
    const uint32_t e = 0x01010101u * row[x < 0 ? 0 : p.W - 1];
    v = make_uint4(e, e, e, e);
  } else {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      w[q] = 0;
#pragma unroll
      for (int i = 0; i < 4; i++) w[q] |= (uint32_t)row[min(x + 4 * q + i, p.W - 1)] << (8 * i);
    }
    v = make_uint4(w[0], w[1], w[2], w[3]);
  }
  *(uint4*)(p.dst - (long)kPad * p.dst_stride - kPad + (long)y * p.dst_stride + x16) = v;
}

__global__ __launch_bounds__(256) void pad_planes_kernel(const PadDesc* __restrict__ pads, int n) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pads[mid].wg_base <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  pad_chunk(pads[lo]);
}

// The same over a handful of planes whose descriptors come as a kernel
// argument (no descriptor upload).
__global__ __launch_bounds__(256) void pad_set_kernel(const PadSet s) {
  PadDesc d = s.d[0];
#pragma unroll
  for (int i = 1; i < kVbsMaxRefs; i++)
    if (i < s.n && s.d[i].wg_base <= (int)blockIdx.x) d = s.d[i];
  pad_chunk(d);
}

// ------------------------------------------------------------- tile bodies

// Tile units (udiv_small's numerators) stay below 16 MBs x 33 quads x 17 row
// tiles.

// Rows are software-pipelined one ahead: the LDS loads of row rho+1 are
// issued before the qsads of row rho.  The empty asm at the top of each row
// takes the row's dwords as operands and clobbers memory: the wait for them
// lands there (a whole row of qsads after their loads were issued), and no
// later row's loads can be hoisted above it (which would exhaust the VGPRs).
// The four 8-byte spans of a row (reference bytes 4k .. 4k+7, k = 0..3) are
// loaded as separate 4-byte-aligned 64-bit values, straight into the VGPR
// pairs v_qsad_pk_u16_u8 takes.
typedef uint64_t __attribute__((aligned(4))) u64_a4;
typedef __attribute__((address_space(3))) u64_a4 lds_u64_a4;

// RDW: the row pitch in dwords when known at compile time (0: row_dw).  With
// it, P rows at a time are read at constant offsets from one base (the
// offsets of ds_read2_b32 reach 255 dwords), so the base moves once per P
// rows instead of once per row.
template <int V, int RDW = 0, typename Done>
__device__ __forceinline__ void sad_tile(const lds_u32* wp, int row_dw,
                                         const uint32_t (&c)[64], Done&& done) {
  constexpr int P = RDW ? (251 / RDW) : 1;
  const int rdw = RDW ? RDW : row_dw;
  uint64_t acc[V];
#pragma unroll
  for (int v = 0; v < V; v++) acc[v] = 0;
  const lds_u32* rp = wp;
  // opaque, so no constant part of wp is folded into the offsets
  if (P > 1) asm volatile("" : "+v"(rp));
  auto span = [&](int rho, int k) { return *(const lds_u64_a4*)(rp + (rho % P) * rdw + k); };
  uint64_t n0 = span(0, 0), n1 = span(0, 1), n2 = span(0, 2), n3 = span(0, 3);
#pragma unroll
  for (int rho = 0; rho < V + 15; rho++) {
    uint64_t p0 = n0, p1 = n1, p2 = n2, p3 = n3;
    // with P == 1 also makes rp opaque, so the next address is one add, not
    // recomputed
    if (P == 1) asm volatile("" : "+v"(p0), "+v"(p1), "+v"(p2), "+v"(p3), "+v"(rp) : : "memory");
    else asm volatile("" : "+v"(p0), "+v"(p1), "+v"(p2), "+v"(p3) : : "memory");
    if (rho + 1 < V + 15) {
      if ((rho + 1) % P == 0) {
        rp += P * rdw;
        if (P > 1) asm volatile("" : "+v"(rp));
      }
      const int nr = rho + 1;
      n0 = span(nr, 0); n1 = span(nr, 1); n2 = span(nr, 2); n3 = span(nr, 3);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint64_t pk = k == 0 ? p0 : (k == 1 ? p1 : (k == 2 ? p2 : p3));
#pragma unroll
      for (int v = 0; v < V; v++) {
        const int r = rho - v;
        if (r >= 0 && r < 16) acc[v] = __builtin_amdgcn_qsad_pk_u16_u8(pk, c[4 * r + k], acc[v]);
      }
    }
    // accumulator rho-15 has seen all 16 block rows: retire it now so its
    // registers free up while the remaining rows run
    if (rho >= 15) done(rho - 15, (uint32_t)acc[rho - 15], (uint32_t)(acc[rho - 15] >> 32));
  }
}

// SSD tile: acc[v][k] = sum over rows of (sum r^2 - 2 sum c*r) for column
// offset k, in wrapping 32-bit arithmetic; the caller adds sum c^2.
template <int V, typename Done>
__device__ __forceinline__ void ssd_tile(const lds_u32* wp, int row_dw,
                                         const uint32_t (&c)[64], Done&& done) {
  uint32_t acc[V][4];
#pragma unroll
  for (int v = 0; v < V; v++)
#pragma unroll
    for (int k = 0; k < 4; k++) acc[v][k] = 0;
  uint32_t nx[5];
  const lds_u32* rp = wp;
#pragma unroll
  for (int i = 0; i < 5; i++) nx[i] = rp[i];
#pragma unroll
  for (int rho = 0; rho < V + 15; rho++) {
    uint32_t d[5];
#pragma unroll
    for (int i = 0; i < 5; i++) d[i] = nx[i];
    // same pipelining as sad_tile
    asm volatile("" : "+v"(d[0]), "+v"(d[1]), "+v"(d[2]), "+v"(d[3]), "+v"(d[4]) : : "memory");
    if (rho + 1 < V + 15) {
      rp += row_dw;
#pragma unroll
      for (int i = 0; i < 5; i++) nx[i] = rp[i];
    }
    // a[kk][k]: reference bytes 4*kk + k .. 4*kk + k + 3 of this row
    uint32_t a[4][4];
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      a[kk][0] = d[kk];
#pragma unroll
      for (int k = 1; k < 4; k++) a[kk][k] = __builtin_amdgcn_alignbyte(d[kk + 1], d[kk], k);
    }
    uint32_t rsq[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      uint32_t s = 0;
#pragma unroll
      for (int kk = 0; kk < 4; kk++) s = __builtin_amdgcn_udot4(a[kk][k], a[kk][k], s, false);
      rsq[k] = s;
    }
#pragma unroll
    for (int v = 0; v < V; v++) {
      const int r = rho - v;
      if (r >= 0 && r < 16) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
          uint32_t cr = 0;
#pragma unroll
          for (int kk = 0; kk < 4; kk++)
            cr = __builtin_amdgcn_udot4(a[kk][k], c[4 * r + kk], cr, false);
          acc[v][k] += rsq[k] - 2u * cr;
        }
      }
    }
    if (rho >= 15) done(rho - 15, acc[rho - 15]);
  }
}

// ------------------------------------------------------------- the kernel
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Lane-wise min of two pairs of packed u16 (v_pk_min_u16).
__device__ __forceinline__ uint32_t min_pk_u16(uint32_t a, uint32_t b) {
  const u16x2 r = __builtin_elementwise_min(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b));
  return __builtin_bit_cast(uint32_t, r);
}

// Fold one candidate into the MB's running key / the cost table.
__device__ __forceinline__ void visit(int dx, int dy, uint32_t cc, int xl, int xr, int yu, int yd,
                                      int L, uint32_t* trow, long long tdim, int* zero_g,
                                      uint64_t& kbest) {
  const bool inwin = dx >= xl && dx <= xr && dy >= yu && dy <= yd;
  if (inwin) {
    uint64_t key;
    if (dx == 0 && dy == 0) {
      *zero_g = (int)cc;
      key = pack_key(cc, 0, kZeroRRank);
    } else {
      key = pack_key(cc, (uint32_t)(dx * dx + dy * dy), raster_rrank(dy, dx));
    }
    kbest = key < kbest ? key : kbest;
  }
  if (trow && dx >= -L && dx <= L && dy >= -L && dy <= L)
    trow[(dy + L) * tdim + (dx + L)] = inwin ? cc : 0xFFFFFFFFu;
}

template <int METRIC, int V, bool TABLE>
__device__ __forceinline__ void run_tile(const StreamDesc& s, const SearchTask& t, int u, int b_lo,
                                         const lds_u32* win, int row_dw, const lds_u32* curs,
                                         const uint32_t* csq, uint64_t* best, int* zero,
                                         uint32_t* __restrict__ table) {
  // unit u: column quad fastest, then MB, then row tile.  Lanes of adjacent
  // MBs read overlapping window dwords (MBs are 4 dwords apart), which the
  // LDS serves as broadcasts instead of bank conflicts.
  const int gq = udiv_small(u, s.Q), q = u - __mul24(gq, s.Q);
  const int bi = udiv_small(gq, t.count), g = gq - __mul24(bi, t.count);
  const int b = b_lo + bi;
  uint32_t c[64];
  const lds_u32x4* cp = (const lds_u32x4*)(curs + g * 64);
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const u32x4 x = cp[i];
    c[4 * i + 0] = x.x;
    c[4 * i + 1] = x.y;
    c[4 * i + 2] = x.z;
    c[4 * i + 3] = x.w;
  }
  const lds_u32* wp = win + (b * V) * row_dw + 4 * g + q;
  const int mb = t.mb + g;
  // a task's MBs lie in one MB row: the division is the task's, not the lane's
  const int row = t.mb / s.mbw, col = t.mb - row * s.mbw + g;
  const int n = col * 16, m = row * 16;
  int xl, xr, yu, yd;
  motion_window(n, m, s.W, s.H, s.L, &xl, &xr, &yu, &yd);
  const int dx0 = tile_dx0(s.L) + 4 * q;
  const int dy0 = -s.L + b * V;
  uint64_t kbest = ~0ull;
  const long long tdim = s.table_dim;
  uint32_t* trow = (TABLE && s.table_base_valid) ? table + s.table_base + (long long)mb * tdim * tdim
                                                 : nullptr;
  if (METRIC == kMetricSAD && TABLE && trow) {
    sad_tile<V>(wp, row_dw, c, [&](int v, uint32_t lo, uint32_t hi) {
      visit(dx0 + 0, dy0 + v, lo & 0xFFFF, xl, xr, yu, yd, s.L, trow, tdim, &zero[g], kbest);
      visit(dx0 + 1, dy0 + v, lo >> 16, xl, xr, yu, yd, s.L, trow, tdim, &zero[g], kbest);
      visit(dx0 + 2, dy0 + v, hi & 0xFFFF, xl, xr, yu, yd, s.L, trow, tdim, &zero[g], kbest);
      visit(dx0 + 3, dy0 + v, hi >> 16, xl, xr, yu, yd, s.L, trow, tdim, &zero[g], kbest);
    });
  } else if (METRIC == kMetricSAD) {
    // Argmin only, branch-free.  Each candidate gets a lane-local 32-bit key
    //   (cost << 16) | code,  code = ((dist2 - d2min) << IB) | (2^IB - 1 - idx)
    // with idx = 4 v + k its raster position in the tile (half) and d2min the
    // least dx^2 + dy^2 over the tile (half): within the lane this orders candidates as
    // the full key does (cost, then distance to the zero predictor, then
    // later raster first).  The host admits a tiling only if every tile's
    // codes fit 16 bits (lane_keys_fit).  code splits into a column part
    // Ck[k] and a row part Ev, so one row costs 3 ORs (window masks), 4 key
    // builds, 3 mins and 2 adds; the full 64-bit key is built once per tile.
    // Candidates outside the window get cost 0xFFFF, above any 16x16 SAD.
    // Tiles taller than 8 rows keep two keys, one per half (rows [0, V1)
    // and [V1, V)), each with its own least dist2: the halves' codes stay
    // within 16 bits at every range up to 64.
    constexpr int V1 = lane_key_split(V);
    constexpr int IB = lane_key_bits(V);
    const uint32_t cm_lo = (dx0 >= xl && dx0 <= xr ? 0u : 0xFFFFu) |
                           (dx0 + 1 >= xl && dx0 + 1 <= xr ? 0u : 0xFFFF0000u);
    const uint32_t cm_hi = (dx0 + 2 >= xl && dx0 + 2 <= xr ? 0u : 0xFFFFu) |
                           (dx0 + 3 >= xl && dx0 + 3 <= xr ? 0u : 0xFFFF0000u);
    const int ax = dx0 > 0 ? dx0 : (dx0 + 3 < 0 ? -(dx0 + 3) : 0);
    // least |dy| over rows [a, a + n)
    auto min_abs = [](int a, int n) { return a > 0 ? a : (a + n - 1 < 0 ? -(a + n - 1) : 0); };
    const int ay = min_abs(dy0, V1);
    uint32_t Ck[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
      Ck[k] = ((uint32_t)((dx0 + k) * (dx0 + k) - ax * ax) << IB) + (uint32_t)(3 - k) + ((1u << IB) - 4u * V1);
    // row part of the code: ((dy^2 - least dy^2 of the half) << IB) +
    // 4 (V1 - 1 - row within the half), affine in v within each half
    const int dyb = dy0 + V1, ayb = min_abs(dyb, V - V1);
    const uint32_t EvA = ((uint32_t)(dy0 * dy0 - ay * ay) << IB) + 4u * (V1 - 1);
    const uint32_t EvB = ((uint32_t)(dyb * dyb - ayb * ayb) << IB) + 4u * (V1 - 1);
    const uint32_t Dl = (uint32_t)(2 * dy0) << IB;
    const uint32_t DlB = (uint32_t)(2 * dyb) << IB;
    // rows of the tile inside [yu, yd]: bit v
    const int vlo = max(yu - dy0, 0), vhi = min(yd - dy0, V - 1);
    const uint32_t inv_rows = vlo > vhi ? 0xFFFFFFFFu : ~(((2u << vhi) - 1u) & ~((1u << vlo) - 1u));
    uint32_t lbest = 0xFFFFFFFFu, lbest0 = 0xFFFFFFFFu;
    uint32_t Evr = EvA;  // the row part, updated row by row (restarting at the second half)
    sad_tile<V>(wp, row_dw, c, [&](int v, uint32_t lo, uint32_t hi) {
      const bool second = v >= V1;
      const uint32_t u = (uint32_t)(second ? v - V1 : v);
      if (V1 < V && v == V1) Evr = EvB;
      const uint32_t Ev = Evr;
      const uint32_t rmask = (uint32_t)__builtin_amdgcn_sbfe((int)inv_rows, v, 1);
      const uint32_t l2 = lo | cm_lo | rmask, h2 = hi | cm_hi | rmask;
      const uint32_t k0 = (l2 << 16) | Ck[0];
      const uint32_t k1 = (l2 & 0xFFFF0000u) | Ck[1];
      const uint32_t k2 = (h2 << 16) | Ck[2];
      const uint32_t k3 = (h2 & 0xFFFF0000u) | Ck[3];
      const uint32_t m = min(min(k0, k1), min(k2, k3)) + Ev;
      if constexpr (V1 < V) {
        if (second) lbest = min(lbest, m);
        else lbest0 = min(lbest0, m);
        // keeps each row's key folding in place (left free, the scheduler
        // hoists it and the tile runs out of registers)
        asm volatile("" : "+v"(lbest), "+v"(lbest0));
      } else {
        lbest0 = min(lbest0, m);
      }
      // Ev of the half's next row: ((dy+1)^2 - dy^2) << IB = (2 dy + 2 u + 1)
      // << IB, minus 4 for the raster part
      Evr += (second ? DlB : Dl) + (((2u * u + 1u) << IB) - 4u);
    });
    auto fold = [&](uint32_t lb, int vbase) {
      if (lb < 0xFFFF0000u) {
        const int idx = (1 << IB) - 1 - (int)(lb & ((1u << IB) - 1u));
        const int dx = dx0 + (idx & 3), dy = dy0 + vbase + (idx >> 2);
        const uint32_t cost = lb >> 16;
        const uint64_t key = (dx == 0 && dy == 0)
                                 ? pack_key(cost, 0, kZeroRRank)
                                 : pack_key(cost, (uint32_t)(dx * dx + dy * dy), raster_rrank(dy, dx));
        if constexpr (V1 < V) kbest = key < kbest ? key : kbest;
        else kbest = key;  // the tile's only key
      }
    };
    fold(lbest0, 0);
    if constexpr (V1 < V) fold(lbest, V1);
  } else {
    const uint32_t c2 = csq[g];
    ssd_tile<V>(wp, row_dw, c, [&](int v, const uint32_t (&a)[4]) {
#pragma unroll
      for (int k = 0; k < 4; k++)
        visit(dx0 + k, dy0 + v, a[k] + c2, xl, xr, yu, yd, s.L, trow, tdim, &zero[g], kbest);
    });
  }
  atomicMin((unsigned long long*)&best[g], (unsigned long long)kbest);
}

// Staging of one task into an LDS buffer with direct global->LDS loads
// (global_load_lds_dwordx4: lane l of a wave writes its 16 bytes at the
// wave's LDS base + 16 l).  The window rows are padded to whole 16-byte
// chunks; the blocks follow the window in the same buffer.
// The window's first column, tile_dx0(L) pels left of the MB, is only
// 4-byte aligned; staging reads from the 16-byte boundary below it and the
// lanes skip the `shift` leading dwords.
__device__ __forceinline__ int stage_shift(int L) { return ((tile_dx0(L) % 16) + 16) % 16 / 4; }

struct Stage {
  int rows, row_dw, row_dw4, win_chunks, cur_chunks, shift;
};

__device__ __forceinline__ Stage stage_shape(const StreamDesc& s, const SearchTask& t) {
  Stage g;
  g.rows = s.B * s.V + 15;
  g.shift = stage_shift(s.L);
  g.row_dw = 4 * (t.count - 1) + s.Q + 4;
  g.row_dw4 = (g.row_dw + g.shift + 3) & ~3;
  g.win_chunks = g.rows * (g.row_dw4 >> 2);
  g.cur_chunks = t.count * 16;
  return g;
}

__device__ __forceinline__ void issue_stage(const StreamDesc& s, const SearchTask& t, lds_u32* buf,
                                            int tid) {
  const Stage g = stage_shape(s, t);
  const int col0 = t.mb % s.mbw, row = t.mb / s.mbw;
  const uint8_t* wsrc =
      s.ref + (long)(16 * row - s.L) * s.ref_stride + 16 * col0 + tile_dx0(s.L) - 4 * g.shift;
  const int cpr = g.row_dw4 >> 2;  // chunks per window row
  const int total = g.win_chunks + g.cur_chunks;
  const int wave = tid >> 6, lane = tid & 63;
  for (int base = wave * 64; base < total; base += kThreads) {
    const int ch = base + lane;
    const uint8_t* src;
    if (ch < g.win_chunks) {
      const int r = ch / cpr, c4 = ch - r * cpr;
      src = wsrc + (long)r * s.ref_stride + 16 * c4;
    } else {
      const int k = ch - g.win_chunks, gm = k >> 4, r = k & 15;
      src = s.cur + (long)(16 * row + r) * s.cur_stride + 16 * (col0 + gm);
    }
    if (ch < total)
      __builtin_amdgcn_global_load_lds((const void*)src,
                                       (__attribute__((address_space(3))) void*)(buf + 4 * base), 16, 0, 0);
  }
}

// The decision of one MB, kept out of line: inlined into the search kernel
// its unrolled loops share one register allocation with the tile loop and
// push the kernel past the VGPR budget of three waves per SIMD.
__device__ __attribute__((noinline)) void decide_in_search(const DecideDesc* d, dec::WaveLds* s, int mb,
                                                           int lane, int mx, int my, int cost) {
  const IntResult r{mx, my, cost, -1};
  dec::decide_mb<true>(*d, *s, mb, 0, 0, lane, &r);
}

// VFIX: tile height fixed at compile time, or 0 to take each stream's own
// (mixed batches).  TABLE: write every cost to the table (filtered scan /
// predictor modes) instead of only the argmin.
// Persistent: workgroup w runs tasks w, w + grid, ... and streams the next
// task's window into the other half of LDS while it computes the current one.
// Waves per SIMD the register allocation must allow: the SAD argmin kernels
// 3 (168 VGPRs; the 8-row one needs 129, one over the 4-wave budget); the
// cost-table kernels keep the compiler's choice.
constexpr int search_min_waves(int metric, int vfix, bool table) {
  return (metric == kMetricSAD && !table) ? 3 : 1;
}

template <int METRIC, int VFIX, bool TABLE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(search_min_waves(METRIC, VFIX, TABLE))))
void me_search_kernel(const StreamDesc* __restrict__ streams, const SearchTask* __restrict__ tasks,
                      int ntasks, IntResult* __restrict__ out, uint32_t* __restrict__ table,
                      int buf_dwords, const DecideDesc* __restrict__ decide, int band) {
  extern __shared__ __attribute__((aligned(16))) uint32_t dyn[];
  __shared__ uint32_t csq[kMaxG];
  __shared__ uint64_t best[kMaxG];
  __shared__ int zero[kMaxG];
  const int tid = threadIdx.x;
  lds_u32* const lds = (lds_u32*)dyn;
  // staging buffer k of the pair
  auto buf = [&](int k) { return lds + (k & 1) * buf_dwords; };

  // Workgroup w runs tasks w, w + grid, w + 2 grid, ... (the host sorts the
  // tasks longest first, so the rounds stay even).  A static schedule: a
  // shared ticket counter costs one device-scope atomic per task on a single
  // address, and those serialize (about 15 ns each, measured), which alone
  // took longer than the search.
  int ti = blockIdx.x;
  // XCD bands: workgroups are dispatched to the 8 XCDs in turn, and each XCD
  // has its own L2.  With one task per workgroup and tasks of equal cost (one
  // stream), XCD x runs a contiguous band of tasks, so neighbouring search
  // windows are fetched from HBM into one L2 instead of into all eight.
  if (band && (int)gridDim.x == ntasks) {
    const int x = (int)blockIdx.x & 7, k = (int)blockIdx.x >> 3;
    const int q = ntasks >> 3, r = ntasks & 7;
    ti = x * q + min(x, r) + k;
  }
  {
    const SearchTask t0 = tasks[ti];
    issue_stage(streams[t0.stream], t0, buf(0), tid);
  }
  if (tid < kMaxG) {
    best[tid] = ~0ull;
    zero[tid] = -1;
  }
  for (int it = 0; ti < ntasks; it++) {
    const SearchTask t = tasks[ti];
    const StreamDesc s = streams[t.stream];
    const Stage g = stage_shape(s, t);
    const lds_u32* win = buf(it) + g.shift;
    const lds_u32* curs = buf(it) + g.win_chunks * 4;
    __builtin_amdgcn_s_waitcnt(0);  // this wave's staging loads have landed
    __syncthreads();                // ... and every other wave's
    if constexpr (!TABLE) __builtin_amdgcn_s_setprio(0);  // tiles at the base priority
    const int tnext = ti + (int)gridDim.x;
    if (tnext < ntasks) {
      const SearchTask tn = tasks[tnext];
      issue_stage(streams[tn.stream], tn, buf(it + 1), tid);
    }
    if (METRIC == kMetricSSD) {
      if (tid < t.count) {
        uint32_t a = 0;
        for (int i = 0; i < 64; i++) a = __builtin_amdgcn_udot4(curs[tid * 64 + i], curs[tid * 64 + i], a, false);
        csq[tid] = a;
      }
      __syncthreads();
    }
    // row tiles that meet the task's window rows (one MB row: the same
    // clamped [yu, yd] for every MB of the task)
    int b_lo, nb;
    {
      const int m = 16 * (t.mb / s.mbw);
      const int yu = (m - s.L >= -16) ? -s.L : -(m + 16);
      const int yd = (m + s.L < s.H) ? s.L : s.H - m;
      b_lo = (yu + s.L) / s.V;
      nb = min((yd + s.L) / s.V, s.B - 1) - b_lo + 1;
    }
    const int units = t.count * s.Q * nb;
    for (int u = tid; u < units; u += kThreads) {
      if constexpr (METRIC == kMetricSSD) {
        run_tile<METRIC, 8, TABLE>(s, t, u, b_lo, win, g.row_dw4, curs, csq, best, zero, table);  // SSD tiles are 8 rows
      } else if constexpr (VFIX != 0) {
        run_tile<METRIC, VFIX, TABLE>(s, t, u, b_lo, win, g.row_dw4, curs, csq, best, zero, table);
      } else {
        switch (s.V) {
          case 8: run_tile<METRIC, 8, TABLE>(s, t, u, b_lo, win, g.row_dw4, curs, csq, best, zero, table); break;
          case 11: run_tile<METRIC, 11, TABLE>(s, t, u, b_lo, win, g.row_dw4, curs, csq, best, zero, table); break;
          case 13: run_tile<METRIC, 13, TABLE>(s, t, u, b_lo, win, g.row_dw4, curs, csq, best, zero, table); break;
          default:
            // 16-row tiles only with cost tables (argmin keys do not fit them
            // beyond small ranges, and their registers would spill here)
            if constexpr (TABLE) run_tile<METRIC, 16, TABLE>(s, t, u, b_lo, win, g.row_dw4, curs, csq, best, zero, table);
            break;
        }
      }
    }
    __syncthreads();
    if constexpr (TABLE) {
      // the decision kernels read the integer results (and the cost table)
      if (tid < t.count) {
        const uint64_t k = best[tid];
        int dx, dy;
        key_vector(k, &dx, &dy);
        IntResult r;
        r.mx = dx;
        r.my = dy;
        r.cost = (int)key_cost(k);
        r.zero = zero[tid];
        out[s.mb_base + t.mb + tid] = r;
        best[tid] = ~0ull;
        zero[tid] = -1;
      }
    } else {
      // Decision in place: wave w finishes MBs w, w + 4, ... of the task
      // (sub-pel refinement, final vector choice, outputs) while the block
      // is still in LDS and the patch is in L2.  Each MB's key is read and
      // reset by the one wave that decides it.
      __shared__ dec::WaveLds ws[kThreads / 64];
      const int wave = tid >> 6, lane = tid & 63;
      // the decision is latency-bound and holds the workgroup's slot: its
      // waves issue ahead of other workgroups' tiles
      __builtin_amdgcn_s_setprio(2);
      for (int gg = wave; gg < t.count; gg += kThreads / 64) {
        const uint64_t k = best[gg];
        int mx, my;
        key_vector(k, &mx, &my);
        dec::wave_sync();
        if (lane == 0) {
          best[gg] = ~0ull;
          zero[gg] = -1;
        }
        decide_in_search(decide + t.stream, ws + wave, t.mb + gg, lane, mx, my, (int)key_cost(k));
      }
    }
    ti = tnext;
  }
}

// ------------------------------------------------------- one MB per wave
// For windows of half-width L = L0 or L0 - 1 with L0 = 8, 16 or 32 a
// workgroup is a single wave that searches one MB.  The MB's block lives in
// SGPRs (every lane of the wave uses the same 64 dwords), which leaves the
// VGPRs to the accumulators, and the wave needs no barrier with any other:
// its own window (2 L0 + 16 rows) is staged with global_load_lds and waited
// for with s_waitcnt.  The 64 lanes tile the 2 L0 x 2 L0 candidates from
// (-L0, -L) as QW column quads x BB row bands of VV rows (16 x 4 x 16 at
// L0 = 32, 8 x 8 x 4 at 16, 4 x 16 x 1 at 8), each band run as tiles of at
// most 8 rows.  L = L0 leaves column +L0 and row +L0: lane l costs
// candidates (L0, l - L0) and (l - L0, L0) with v_sad_u8 against LDS rows
// (the row's reference bytes start at byte l, realigned with alignbyte),
// and candidate (L0, L0) is one block dword per lane summed over the wave.
// Each tile folds its candidates into one lane-local key as run_tile does;
// the wave's minimum key is then the reference's decision, and the same wave
// decides the MB (sub-pel, outputs) with the window's LDS reused.

// One lane tile of R candidate rows x 4 columns from (dx0, dy0): the least
// lane key (cost << 16) | ((dist2 - least dist2 of the tile) << IB) |
// (2^IB - 1 - raster index), folded into the 64-bit key.  The codes fit 16
// bits for tiles of up to 8 rows at every range up to 64 (IB = 5).
// MASKED: some candidate of the pass may lie outside the MB's window (the
// caller knows it per wave); without it the per-row masks are left out.
template <int R, int IB, int ROW_DW, bool MASKED>
__device__ __forceinline__ void wave_tile(const lds_u32* wp, const uint32_t (&c)[64], int dx0, int dy0, int xl,
                                          int xr, int yu, int yd, uint64_t& kbest) {
  const uint32_t cm_lo = !MASKED ? 0u
                                 : (dx0 >= xl && dx0 <= xr ? 0u : 0xFFFFu) |
                                       (dx0 + 1 >= xl && dx0 + 1 <= xr ? 0u : 0xFFFF0000u);
  const uint32_t cm_hi = !MASKED ? 0u
                                 : (dx0 + 2 >= xl && dx0 + 2 <= xr ? 0u : 0xFFFFu) |
                                       (dx0 + 3 >= xl && dx0 + 3 <= xr ? 0u : 0xFFFF0000u);
  const int ax = dx0 > 0 ? dx0 : (dx0 + 3 < 0 ? -(dx0 + 3) : 0);
  const int ay = dy0 > 0 ? dy0 : (dy0 + R - 1 < 0 ? -(dy0 + R - 1) : 0);
  uint32_t Ck[4];
#pragma unroll
  for (int k = 0; k < 4; k++)
    Ck[k] = ((uint32_t)((dx0 + k) * (dx0 + k) - ax * ax) << IB) + (uint32_t)(3 - k) + ((1u << IB) - 4u * R);
  uint32_t Ev = ((uint32_t)(dy0 * dy0 - ay * ay) << IB) + 4u * (R - 1);
  const uint32_t Dl = (uint32_t)(2 * dy0) << IB;
  const int vlo = max(yu - dy0, 0), vhi = min(yd - dy0, R - 1);
  const uint32_t inv_rows = vlo > vhi ? 0xFFFFFFFFu : ~(((2u << vhi) - 1u) & ~((1u << vlo) - 1u));
  uint32_t lbest = 0xFFFFFFFFu;
  sad_tile<R, ROW_DW>(wp, ROW_DW, c, [&](int v, uint32_t lo, uint32_t hi) {
    const uint32_t rmask = MASKED ? (uint32_t)__builtin_amdgcn_sbfe((int)inv_rows, v, 1) : 0u;
    const uint32_t l2 = MASKED ? lo | cm_lo | rmask : lo, h2 = MASKED ? hi | cm_hi | rmask : hi;
    const uint32_t k0 = (l2 << 16) | Ck[0];
    const uint32_t k1 = (l2 & 0xFFFF0000u) | Ck[1];
    const uint32_t k2 = (h2 << 16) | Ck[2];
    const uint32_t k3 = (h2 & 0xFFFF0000u) | Ck[3];
    lbest = min(lbest, min(min(k0, k1), min(k2, k3)) + Ev);
    // Ev of the next row: ((dy+1)^2 - dy^2) << IB, minus 4 for the raster part
    Ev += Dl + (((2u * (uint32_t)v + 1u) << IB) - 4u);
  });
  if (lbest < 0xFFFF0000u) {
    const int idx = (1 << IB) - 1 - (int)(lbest & ((1u << IB) - 1u));
    const int dx = dx0 + (idx & 3), dy = dy0 + (idx >> 2);
    const uint32_t cost = lbest >> 16;
    const uint64_t key = (dx == 0 && dy == 0) ? pack_key(cost, 0, kZeroRRank)
                                              : pack_key(cost, (uint32_t)(dx * dx + dy * dy), raster_rrank(dy, dx));
    kbest = key < kbest ? key : kbest;
  }
}

__device__ __forceinline__ void fold_key(uint64_t& kbest, bool live, uint32_t cost, int dx, int dy, int xl, int xr,
                                         int yu, int yd) {
  if (live && dx >= xl && dx <= xr && dy >= yu && dy <= yd) {
    const uint64_t key = pack_key(cost, (uint32_t)(dx * dx + dy * dy), raster_rrank(dy, dx));
    kbest = key < kbest ? key : kbest;
  }
}

// Geometry of one window class.
template <int L0>
struct WaveGeom {
  static constexpr int QW = L0 / 2;                        // column quads: 2 L0 columns
  static constexpr int BB = 64 / QW;                       // row bands
  static constexpr int VV = 2 * L0 / BB;                   // rows per band
  static constexpr int TR = VV < 8 ? VV : 8;               // rows per tile
  static constexpr int SHIFT = ((16 - L0 % 16) % 16) / 4;  // dwords from the 16-byte boundary to column -L0
  static constexpr int ROW_DW = ((2 * L0 + 16 + 4 * SHIFT + 15) / 16) * 4;
  static constexpr int ROWS = 2 * L0 + 16;
  static constexpr int WIN_BYTES = ROWS * ROW_DW * 4;
  static constexpr int LDS_BYTES =
      WIN_BYTES > (int)sizeof(dec::WaveLds) ? WIN_BYTES : (int)((sizeof(dec::WaveLds) + 15) / 16 * 16);
  static_assert(QW * BB == 64 && BB * VV == 2 * L0 && VV % TR == 0, "lane tiling");
};

// The MB a wave searches: its position and window.
struct WaveMb {
  int n, m, L;
  int xl, xr, yu, yd;
};

__device__ __forceinline__ WaveMb wave_mb(const StreamDesc& s, int mb) {
  WaveMb w;
  const int row = mb / s.mbw, col = mb - row * s.mbw;
  w.n = 16 * col;
  w.m = 16 * row;
  w.L = s.L;
  motion_window(w.n, w.m, s.W, s.H, w.L, &w.xl, &w.xr, &w.yu, &w.yd);
  return w;
}

// One pass of the window class L0 over the 2 L0 x 2 L0 candidates whose
// top-left is (dxb, dyb) (dxb - L0 a multiple of 16), folded into kbest:
// the window rows m + dyb .. m + dyb + 2 L0 + 15 are staged from the
// 16-byte boundary at or left of column n + dxb, then every lane runs its
// tiles.  lcol / lrow: also the column dx = dxb + 2 L0 / the row
// dy = dyb + 2 L0 (64 candidates each, lane l the l-th), and corner the
// candidate where they meet.  Candidates outside the MB's window fold
// nothing.  The caller has finished every LDS read of an earlier pass.
template <int L0>
__device__ __forceinline__ void wave_pass(const StreamDesc& s, const WaveMb& w, lds_u32* win, int lane, int dxb,
                                          int dyb, bool lcol, bool lrow, bool corner, uint64_t& kbest) {
  using G = WaveGeom<L0>;
  // the block, loaded by every pass (nothing lives across passes): uniform
  // addresses, so scalar loads straight into SGPRs (through the constant
  // address space, where uniform loads become s_load)
  uint32_t c[64];
  {
    typedef __attribute__((address_space(4))) const uint32_t const_u32;
    const const_u32* cb = (const const_u32*)(s.cur + (long)w.m * s.cur_stride + w.n);
    const int sdw = s.cur_stride >> 2;
#pragma unroll
    for (int i = 0; i < 16; i++)
#pragma unroll
      for (int k = 0; k < 4; k++) c[4 * i + k] = cb[i * sdw + k];
  }
  {
    // chunk base + lane of the window: row r, 16-byte column c4, at a 32-bit
    // offset from the uniform row origin (a scalar base and one VGPR offset
    // per load)
    constexpr int kCpr = G::ROW_DW / 4, kChunks = G::ROWS * kCpr;
    const uint8_t* wsrc = s.ref + (long)(w.m + dyb) * s.ref_stride + w.n + dxb - 4 * G::SHIFT;
    __builtin_assume(lane >= 0 && lane < 64);
#pragma unroll
    for (int base = 0; base < kChunks; base += 64) {
      // r * stride + 16 (ch - kCpr r), strides below 2^23
      const uint32_t ch = (uint32_t)(base + lane);
      const uint32_t r = ch / kCpr;
      const uint32_t off = (uint32_t)__mul24((int)r, s.ref_stride - 16 * kCpr) + (ch << 4);
      if (base + 64 <= kChunks || ch < (uint32_t)kChunks)
        __builtin_amdgcn_global_load_lds((const void*)(wsrc + off),
                                         (__attribute__((address_space(3))) void*)(win + 4 * base), 16, 0, 0);
    }
  }
  __builtin_amdgcn_s_waitcnt(0);  // the block and the window have landed (one wave: no barrier)
  dec::wave_sync();

  const lds_u32* w0 = win + G::SHIFT;  // column dxb
  const int q = lane % G::QW, b = lane / G::QW;
  const int dx0 = dxb + 4 * q;
  const int dy0 = dyb + b * G::VV;
  // every candidate of the pass inside the window (the interior MBs): the
  // tiles need no masks
  const bool inside = dxb >= w.xl && dxb + 2 * L0 - 1 <= w.xr && dyb >= w.yu && dyb + 2 * L0 - 1 <= w.yd;
  if (inside) {
#pragma unroll
    for (int h = 0; h < G::VV / G::TR; h++)
      wave_tile<G::TR, 5, G::ROW_DW, false>(w0 + (b * G::VV + h * G::TR) * G::ROW_DW + q, c, dx0, dy0 + h * G::TR,
                                            w.xl, w.xr, w.yu, w.yd, kbest);
  } else {
#pragma unroll
    for (int h = 0; h < G::VV / G::TR; h++)
      wave_tile<G::TR, 5, G::ROW_DW, true>(w0 + (b * G::VV + h * G::TR) * G::ROW_DW + q, c, dx0, dy0 + h * G::TR,
                                           w.xl, w.xr, w.yu, w.yd, kbest);
  }
  const bool live = lane < 2 * L0;
  const int l = live ? lane : 2 * L0 - 1;  // lanes past the line read in bounds, fold nothing
  if (lcol) {
    // candidate (dxb + 2 L0, dyb + l) from window rows l .. l + 15, bytes
    // 2 L0 .. 2 L0 + 15
    uint32_t acc = 0;
    const lds_u32* rp = w0 + l * G::ROW_DW + L0 / 2;
#pragma unroll
    for (int r = 0; r < 16; r++)
#pragma unroll
      for (int k = 0; k < 4; k++) acc = __builtin_amdgcn_sad_u8(rp[r * G::ROW_DW + k], c[4 * r + k], acc);
    fold_key(kbest, live, acc, dxb + 2 * L0, dyb + l, w.xl, w.xr, w.yu, w.yd);
  }
  if (lrow) {
    // candidate (dxb + l, dyb + 2 L0), l = 4 g + t, from window rows
    // 2 L0 .. 2 L0 + 15: lane l sums block rows 4 t .. 4 t + 3 (loaded by
    // the lane itself) against candidates 4 g .. 4 g + 3, one
    // v_qsad_pk_u16_u8 per block dword on dword-aligned reference bytes;
    // the quad's four partial sums (each below 2^16 per candidate, so two
    // 16-bit sums share a dword) are added by DPP and lane l keeps its own.
    const int g = (live ? lane : 2 * L0 - 1) >> 2, t = lane & 3;
    const lds_u32* rp = w0 + (2 * L0 + 4 * t) * G::ROW_DW + g;
    const uint32_t* cb = (const uint32_t*)(s.cur + (long)(w.m + 4 * t) * s.cur_stride + w.n);
    const int sdw = s.cur_stride >> 2;
    uint64_t acc = 0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      uint32_t v[5];
#pragma unroll
      for (int j = 0; j < 5; j++) v[j] = rp[r * G::ROW_DW + j];
#pragma unroll
      for (int k = 0; k < 4; k++)
        acc = __builtin_amdgcn_qsad_pk_u16_u8(((uint64_t)v[k + 1] << 32) | v[k], cb[r * sdw + k], acc);
    }
    uint32_t lo = (uint32_t)acc, hi = (uint32_t)(acc >> 32);
    lo += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0xB1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
    hi += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0xB1, 0xf, 0xf, false);
    lo += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0x4E, 0xf, 0xf, false);  // quad_perm [2,3,0,1]
    hi += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0x4E, 0xf, 0xf, false);
    const uint32_t cost = ((t & 2 ? hi : lo) >> (16 * (t & 1))) & 0xFFFFu;
    fold_key(kbest, live, cost, dxb + l, dyb + 2 * L0, w.xl, w.xr, w.yu, w.yd);
  }
  if (corner) {
    // lane l the block dword (l / 4, l % 4)
    const uint32_t cw = *(const uint32_t*)(s.cur + (long)(w.m + (lane >> 2)) * s.cur_stride + w.n + 4 * (lane & 3));
    const uint32_t rw = w0[(2 * L0 + (lane >> 2)) * G::ROW_DW + L0 / 2 + (lane & 3)];
    const int tot = dec::wave_sum((int)__builtin_amdgcn_sad_u8(cw, rw, 0u));
    fold_key(kbest, true, (uint32_t)tot, dxb + 2 * L0, dyb + 2 * L0, w.xl, w.xr, w.yu, w.yd);
  }
}

// The integer search of one MB whose window class is L0 (8, 16, 32 or 64):
// one pass, or for L0 = 64 four quadrant passes of the 32 class (candidate
// columns from -64 and from 0, rows from -L and from 64 - L); quadrants
// wholly outside the MB's window are skipped.  Returns the MB's least key.
__device__ __forceinline__ uint64_t wave_search_mb(const StreamDesc& s, const WaveMb& w, lds_u32* win, int lane,
                                                   int L0) {
  uint64_t kbest = ~0ull;
  const bool ex = w.L == L0;  // exhaustive: one more column and row
  if (L0 == 8) {
    wave_pass<8>(s, w, win, lane, -8, -w.L, ex, ex, ex, kbest);
  } else if (L0 == 16) {
    wave_pass<16>(s, w, win, lane, -16, -w.L, ex, ex, ex, kbest);
  } else if (L0 == 32) {
    wave_pass<32>(s, w, win, lane, -32, -w.L, ex, ex, ex, kbest);
  } else {
    const int np = 4;
#pragma unroll 1
    for (int p = 0; p < np; p++) {
      const int qx = p & 1, qy = p >> 1;
      const int dxb = np == 1 ? -32 : -64 + 64 * qx, dyb = -w.L + 64 * qy;
      const bool lc = ex && (np == 1 || qx), lr = ex && (np == 1 || qy);
      const int dxe = dxb + 63 + (lc ? 1 : 0), dye = dyb + 63 + (lr ? 1 : 0);
      if (dxe < w.xl || dxb > w.xr || dye < w.yu || dyb > w.yd) continue;
      dec::wave_sync();  // an earlier pass's LDS reads are done
      int ln = lane;  // opaque per pass, so no per-lane address is held across passes
      asm volatile("" : "+v"(ln));
      __builtin_assume(ln >= 0 && ln < 64);
      wave_pass<32>(s, w, win, ln, dxb, dyb, lc, lr, lc && lr, kbest);
    }
  }
  return dec::wave_min_u64_uniform(kbest);
}

// The decision of a searched MB, in the same wave, with the window's LDS
// reused (at a raised wave priority: it is the MB's serial tail).  LEAN: the
// launch's streams are all of the lean modes (SAD, integer or half pel, no
// final vector choice).
template <bool LEAN>
__device__ __forceinline__ void wave_decide_mb(const DecideDesc& d, uint32_t* lds, int mb, int lane,
                                               uint64_t kbest) {
  int mx, my;
  key_vector(kbest, &mx, &my);
  dec::wave_sync();  // every lane's window reads are done: the decision reuses the LDS
  __builtin_amdgcn_s_setprio(2);
  // inlined: with the block in SGPRs the tile and the decision fit 80 VGPRs
  // together (6 waves per SIMD); called out of line the decision alone took 87
  const IntResult r{mx, my, (int)key_cost(kbest), -1};
  if (LEAN) dec::decide_mb_lean(d, mb, lane, r);
  else dec::decide_mb<true>(d, *(dec::WaveLds*)lds, mb, 0, 0, lane, &r);
}

// LDS of any window class (the 64 class stages 32-class quadrants).
constexpr int kWaveLdsBytes =
    WaveGeom<32>::LDS_BYTES > WaveGeom<16>::LDS_BYTES
        ? (WaveGeom<32>::LDS_BYTES > WaveGeom<8>::LDS_BYTES ? WaveGeom<32>::LDS_BYTES : WaveGeom<8>::LDS_BYTES)
        : (WaveGeom<16>::LDS_BYTES > WaveGeom<8>::LDS_BYTES ? WaveGeom<16>::LDS_BYTES : WaveGeom<8>::LDS_BYTES);

template <int L0, bool LEAN>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6)))
void me_wave_kernel(const StreamDesc* __restrict__ streams, const SearchTask* __restrict__ tasks, int ntasks,
                    const DecideDesc* __restrict__ decide, int band) {
  constexpr int kLds = L0 == 64 ? WaveGeom<32>::LDS_BYTES : WaveGeom<L0>::LDS_BYTES;
  __shared__ __attribute__((aligned(16))) uint32_t win_raw[kLds / 4];
  const int lane = threadIdx.x;
  int ti = blockIdx.x;
  if (band) {
    // XCD bands, as in me_search_kernel
    const int x = (int)blockIdx.x & 7, k = (int)blockIdx.x >> 3;
    const int q = ntasks >> 3, r = ntasks & 7;
    ti = x * q + min(x, r) + k;
  }
  const SearchTask t = tasks[ti];
  const StreamDesc s = streams[t.stream];
  const uint64_t kbest = wave_search_mb(s, wave_mb(s, t.mb), (lds_u32*)win_raw, lane, L0);
  wave_decide_mb<LEAN>(decide[t.stream], win_raw, t.mb, lane, kbest);
}

// Persistent form over a macroblock work queue: a grid of one-wave
// workgroups, as many as stay resident, in which wave b runs the tasks
// [woff[b], woff[b + 1]) of `tasks` (one MB each, any window class) one
// after another.  The host builds the queue: the MBs are split over the 8
// XCDs (workgroup b runs on XCD b mod 8) in contiguous runs per window
// class, and each XCD's share is dealt to its waves costliest first, each MB
// to the wave with the least work so far (LPT), by an estimate of the MB's
// work; a wave's list is costliest first too, so waves of an XCD run
// neighbouring MBs side by side and every wave ends with small ones.  No
// atomics: a queue taken from a shared counter measured twice the time of
// the one-launch-per-class plan on a single 1080p stream, its counter being
// contended by every wave of an XCD.
// BIG: the +-31/32 and +-63/64 classes only, so the kernel carries one
// search path (the 32-class pass) beside the decision.
template <bool BIG, bool LEAN>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6)))
void me_queue_kernel(const StreamDesc* __restrict__ streams, const SearchTask* __restrict__ tasks,
                     const int* __restrict__ woff, const DecideDesc* __restrict__ decide) {
  __shared__ __attribute__((aligned(16))) uint32_t win_raw[kWaveLdsBytes / 4];
  const int lane = threadIdx.x;
  const int beg = woff[blockIdx.x], end = woff[blockIdx.x + 1];
  for (int ti = beg; ti < end; ti++) {
    __builtin_amdgcn_s_setprio(0);
    // the lane index made opaque per task: otherwise every per-lane address
    // of the task would be hoisted out of the loop and held (spilled) across it
    int ln = lane;
    asm volatile("" : "+v"(ln));
    __builtin_assume(ln >= 0 && ln < 64);
    const SearchTask t = tasks[ti];
    const StreamDesc s = streams[t.stream];
    const int L0 = BIG ? (s.L <= 32 ? 32 : 64) : (s.L <= 8 ? 8 : (s.L <= 16 ? 16 : (s.L <= 32 ? 32 : 64)));
    const uint64_t kbest = wave_search_mb(s, wave_mb(s, t.mb), (lds_u32*)win_raw, ln, L0);
    wave_decide_mb<LEAN>(decide[t.stream], win_raw, t.mb, ln, kbest);
    dec::wave_sync();  // the decision's LDS reads are done before the next staging
  }
}

void launch_wave_search(const StreamDesc* d_streams, const SearchTask* d_tasks, int ntasks, int L0,
                        const DecideDesc* d_decide, bool band, bool lean, hipStream_t st) {
  if (ntasks <= 0) return;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(ntasks), dim3(64), 0, st, d_streams, d_tasks, ntasks, d_decide, band ? 1 : 0);
  };
  if (lean) {
    if (L0 == 64) go(me_wave_kernel<64, true>);
    else if (L0 == 32) go(me_wave_kernel<32, true>);
    else if (L0 == 16) go(me_wave_kernel<16, true>);
    else go(me_wave_kernel<8, true>);
  } else {
    if (L0 == 64) go(me_wave_kernel<64, false>);
    else if (L0 == 32) go(me_wave_kernel<32, false>);
    else if (L0 == 16) go(me_wave_kernel<16, false>);
    else go(me_wave_kernel<8, false>);
  }
}

void launch_wave_queue(const StreamDesc* d_streams, const SearchTask* d_tasks, const int* d_woff, int waves,
                       bool big_only, bool lean, const DecideDesc* d_decide, hipStream_t st) {
  if (waves <= 0) return;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(waves), dim3(64), 0, st, d_streams, d_tasks, d_woff, d_decide);
  };
  if (big_only) {
    if (lean) go(me_queue_kernel<true, true>);
    else go(me_queue_kernel<true, false>);
  } else {
    if (lean) go(me_queue_kernel<false, true>);
    else go(me_queue_kernel<false, false>);
  }
}

long wave_task_cost(int n, int m, int W, int H, int L) {
  // In candidates: per pass the 2 L0 x 2 L0 grid (and the exhaustive
  // window's extra column and row) plus 780 for staging the window, plus 130
  // for the decision.  Fitted to the one-launch-per-class times of a wave
  // slot at 6 waves per SIMD: 13.0 / 21.8 / 56 / 217 us per MB at
  // +-8 / 16 / 32 / 64, about 0.0109 us per candidate.
  const int L0 = wave_class(L);
  const long ex = L == L0 ? 1 : 0;
  if (L0 != 64) return 4L * L0 * L0 + ex * (4L * L0 + 1) + 780 + 130;
  int xl, xr, yu, yd;
  motion_window(n, m, W, H, L, &xl, &xr, &yu, &yd);
  long c = 130;
  for (int p = 0; p < 4; p++) {
    const int qx = p & 1, qy = p >> 1;
    const int dxb = -64 + 64 * qx, dyb = -L + 64 * qy;
    const int dxe = dxb + 63 + (ex && qx ? 1 : 0), dye = dyb + 63 + (ex && qy ? 1 : 0);
    if (dxe < xl || dxb > xr || dye < yu || dyb > yd) continue;
    c += 4096 + 64 * ex * (qx + qy) + 780;
  }
  return c;
}

int wave_class(int L) {
  for (int L0 : {8, 16, 32, 64})
    if (L == L0 || L == L0 - 1) return L0;
  return 0;
}

// ----------------------------------------------------------------- host side
int pad_plane_workgroups(int W, int H) {
  const long chunks = (long)(W + 2 * kPad) / 16 * (H + 2 * kPad);
  return (int)((chunks + 255) / 256);
}

void launch_pad_set(const PadSet& s, int total_wgs, hipStream_t st) {
  if (s.n <= 0 || total_wgs <= 0) return;
  hipLaunchKernelGGL(pad_set_kernel, dim3(total_wgs), dim3(256), 0, st, s);
}

void launch_pad_planes(const PadDesc* d_pads, int n, int total_wgs, hipStream_t st) {
  if (n <= 0 || total_wgs <= 0) return;
  hipLaunchKernelGGL(pad_planes_kernel, dim3(total_wgs), dim3(256), 0, st, d_pads, n);
}

// vfix: the common tile height of every task, or 0 when streams differ.
// buf_dwords: one staging buffer (window + blocks) of the largest task; the
// kernel holds two.  table: some stream keeps a cost table; then the
// decision kernels run after the search, otherwise the search kernel decides
// each MB itself from d_decide.
// grid_mode: 0 one task per workgroup; 1 persistent, as many workgroups as
// stay resident; 2 persistent, that many reduced so that every workgroup runs
// the same number of tasks (no partly filled last round).
void launch_search(const StreamDesc* d_streams, const SearchTask* d_tasks, int ntasks, int grid_mode,
                   int metric, int vfix, bool table, int buf_dwords, IntResult* d_out,
                   uint32_t* d_table, const DecideDesc* d_decide, bool band, hipStream_t st) {
  const size_t shmem = 2 * (size_t)buf_dwords * sizeof(uint32_t);
  auto go = [&](auto kern) {
    int grid = ntasks;
    if (grid_mode != 0) {
      // as many workgroups as stay resident at once (registers and LDS),
      // remembered per kernel and LDS size
      static thread_local const void* last_kern = nullptr;
      static thread_local size_t last_shmem = 0;
      static thread_local int last_dev = -1, last_res = 0;
      int dev = 0;
      hipGetDevice(&dev);
      if (last_kern != (const void*)kern || last_shmem != shmem || last_dev != dev) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kThreads, shmem) != hipSuccess)
          per_cu = 1;
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        last_kern = (const void*)kern;
        last_shmem = shmem;
        last_dev = dev;
        last_res = std::max(1, per_cu) * std::max(1, cus);
      }
      grid = std::min(ntasks, last_res);
      if (grid_mode == 2) {
        const int rounds = (ntasks + grid - 1) / grid;
        grid = (ntasks + rounds - 1) / rounds;
      }
    }
    // a workgroup that runs one task never touches the second buffer
    const size_t lds_bytes = grid >= ntasks ? shmem / 2 : shmem;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), lds_bytes, st, d_streams, d_tasks, ntasks, d_out,
                       d_table, buf_dwords, d_decide, band ? 1 : 0);
  };
  if (metric == kMetricSSD) {
    if (table) go(me_search_kernel<kMetricSSD, 8, true>);
    else go(me_search_kernel<kMetricSSD, 8, false>);
    return;
  }
  if (table) {
    switch (vfix) {
      case 8: go(me_search_kernel<kMetricSAD, 8, true>); break;
      case 11: go(me_search_kernel<kMetricSAD, 11, true>); break;
      case 13: go(me_search_kernel<kMetricSAD, 13, true>); break;
      case 16: go(me_search_kernel<kMetricSAD, 16, true>); break;
      default: go(me_search_kernel<kMetricSAD, 0, true>); break;
    }
  } else {
    switch (vfix) {
      case 8: go(me_search_kernel<kMetricSAD, 8, false>); break;
      case 11: go(me_search_kernel<kMetricSAD, 11, false>); break;
      case 13: go(me_search_kernel<kMetricSAD, 13, false>); break;
      default: go(me_search_kernel<kMetricSAD, 0, false>); break;
    }
  }
}

int search_max_window_dwords() { return kWinDwords; }
int search_stage_dwords(int B, int V, int Q, int G) {
  const int rows = B * V + 15, row_dw4 = (4 * (G - 1) + Q + 4 + 3 + 3) & ~3;  // + largest shift
  return rows * row_dw4 + 64 * G;
}
int search_max_group() { return kMaxG; }

}  // namespace icu
