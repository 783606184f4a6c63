// Multi-reference, variable-block-size integer motion search on gfx950.
//
// Per 16x16 MB and per reference frame, every full-pel candidate of the MB's
// window (GetMotionRange with half-width R,
// src/CodecUtils/MotionEstimatorH264ImplFull.cpp:1029-1081, exhaustive form)
// gets the SADs of the MB's four 8x8 quadrants, each the sum of four 4x4
// SADs (OverlayMem2Dv2::Tad4x4 / Tad8x8, src/ImageUtils/OverlayMem2Dv2.cpp:
// 2555-2654).  The quadrant SADs are merged into the nine partitions of the
// 16x16 / 16x8 / 8x16 / 8x8 modes, a rate term lambda * (|vx| + |vy|) is
// added (no reference counterpart: the reference searches 16x16 blocks of
// one reference), and each partition keeps its least cost; ties go to the
// lower reference index, then to the earlier candidate in raster order.  The
// MB then takes the mode whose partitions cost least in total, ties to the
// mode with fewer partitions.
//
// Work decomposition: one workgroup per MB.  The windows of all references
// (2R+1 candidate rows + 15, dword rows) are staged in LDS with
// global_load_lds; the MB's block is read once with scalar loads (every lane
// of the workgroup uses the same 64 dwords, so they live in SGPRs).  A lane
// task is one reference x one 4-candidate column quad x V candidate rows: it
// walks the V+15 reference rows its tile touches once, and each (reference
// row, block row) pair issues 4 v_qsad_pk_u16_u8 into the accumulators of
// the quadrant that block row and column half belong to (an 8x8 SAD is at
// most 16320, so the packed u16 sums never wrap).  Each finished candidate
// row merges the quadrants into the nine partitions and folds them into nine
// lane-local minimum keys (cost << 15 | tie-break index); the workgroup then
// reduces the keys across lanes and waves and decides the mode.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "icu_device.h"
#include "icu_internal.h"

namespace icu {

namespace {

typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef uint64_t __attribute__((aligned(4))) u64_a4;
typedef __attribute__((address_space(3))) u64_a4 lds_u64_a4;

constexpr int kThreads = 256;

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Two u16 SADs in one dword added lane-wise (v_pk_add_u16): partition sums
// of up to four 8x8 SADs stay below 65536.
__device__ __forceinline__ uint32_t add_u16x2(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) + __builtin_bit_cast(u16x2, b));
}

// A candidate's key, half * 2^15 + off, from one u16 half of a packed SAD
// pair in a single v_mad_u32_u16 (op_sel picks the high half); k15 holds
// 2^15 in a register (VOP3 takes no literal here; the SGPRs hold the block).
__device__ __forceinline__ uint32_t key_lo16(uint32_t x, uint32_t k15, uint32_t off) {
  uint32_t r;
  asm("v_mad_u32_u16 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(k15), "v"(off));
  return r;
}
__device__ __forceinline__ uint32_t key_hi16(uint32_t x, uint32_t k15, uint32_t off) {
  uint32_t r;
  asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[1,0,0,0]" : "=v"(r) : "v"(x), "v"(k15), "v"(off));
  return r;
}

// Partition order: 0 16x16; 1, 2 16x8 top / bottom; 3, 4 8x16 left / right;
// 5..8 8x8 in raster order.  Quadrants: 0 top-left, 1 top-right,
// 2 bottom-left, 3 bottom-right.
__device__ __forceinline__ uint32_t min3_u32(uint32_t a, uint32_t b, uint32_t c) { return min(min(a, b), c); }

__device__ __forceinline__ int part_of_quadrant(int mode, int k) {
  return mode == 0 ? 0 : (mode == 1 ? 1 + (k >> 1) : (mode == 2 ? 3 + (k & 1) : 5 + k));
}

// Waves per SIMD each tile height is compiled for: its accumulators (4 x V
// packed pairs) and the epilogue fit without spilling.
constexpr int vbs_max_waves(int V) { return V <= 4 ? 5 : (V <= 7 ? 4 : 3); }

template <int V>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, vbs_max_waves(V))))
void vbs_search_kernel(const VbsParams p) {
  extern __shared__ __attribute__((aligned(16))) uint32_t dyn[];
  __shared__ uint32_t red[kThreads / 64][9];
  lds_u32* const win = (lds_u32*)dyn;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  // XCD bands: workgroups go to the 8 XCDs in turn; XCD x runs a contiguous
  // band of MBs, so neighbouring windows are fetched into one L2.
  int mb;
  {
    const int x = (int)blockIdx.x & 7, k = (int)blockIdx.x >> 3;
    const int q = p.nmb >> 3, r = p.nmb & 7;
    mb = x * q + min(x, r) + k;
  }
  const int mcol = mb % p.mbw, mrow = mb / p.mbw;
  const int n = 16 * mcol, m = 16 * mrow;

  // stage every reference's window: rows m-R .. m-R+rows-1, 16-byte chunks
  // from the 16-byte boundary at or left of column n + dx0 (chunk counts
  // stay within udiv_small's domain: vbs_plan checks them)
  {
    const int cpr = p.row_dw >> 2;
    const int per_ref = p.rows * cpr;
    const int total = p.nrefs * per_ref;
    for (int base = wave * 64; base < total; base += kThreads) {
      const int ch = base + lane;
      if (ch < total) {
        const int r = udiv_small(ch, per_ref), rem = ch - r * per_ref;
        const int row = udiv_small(rem, cpr), c4 = rem - row * cpr;
        const uint8_t* rb = r == 0 ? p.ref[0] : (r == 1 ? p.ref[1] : (r == 2 ? p.ref[2] : p.ref[3]));
        const uint8_t* src = rb + (long)(m - p.R + row) * p.ref_stride + n + p.dx0 - 4 * p.shift + 16 * c4;
        __builtin_amdgcn_global_load_lds((const void*)src,
                                         (__attribute__((address_space(3))) void*)(win + 4 * base), 16, 0, 0);
      }
    }
  }
  // the MB's block: uniform addresses, so scalar loads into SGPRs
  uint32_t c[64];
  {
    const uint32_t* cw = (const uint32_t*)(p.cur + (long)m * p.W + n);
    const int sdw = p.W >> 2;
#pragma unroll
    for (int i = 0; i < 16; i++)
#pragma unroll
      for (int k = 0; k < 4; k++) c[4 * i + k] = __builtin_amdgcn_readfirstlane(cw[i * sdw + k]);
  }
  int xl, xr, yu, yd;
  motion_window(n, m, p.W, p.H, p.R, &xl, &xr, &yu, &yd);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  uint32_t best[9];
#pragma unroll
  for (int i = 0; i < 9; i++) best[i] = ~0u;
  const int QG = p.Q * p.Gr;
  const int T = p.nrefs * QG;
  const int NN = p.N * p.N;
  const uint32_t k15 = 1u << 15;
  for (int t = tid; t < T; t += kThreads) {
    const int r = udiv_small(t, QG), rem = t - r * QG;
    const int g = udiv_small(rem, p.Q), q = rem - g * p.Q;
    const lds_u32* rp = win + (r * p.rows + g * V) * p.row_dw + p.shift + q;
    const int dx0 = p.dx0 + 4 * q, dy0 = -p.R + g * V;
    // Key offsets (rate << 15) + tie-break index, split into a column part
    // (fixed per task) and a row part (per candidate row).  Candidates
    // outside the MB's window get the offset 2^31 instead: keys stay below
    // 2^32 (cost < 2^17), and the window always holds the zero vector, whose
    // key is below 2^31, so they never win.
    uint32_t offc[4];
    bool colok[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int dx = dx0 + j;
      colok[j] = dx >= xl && dx <= xr;
      offc[j] = ((uint32_t)(p.lambda * (dx < 0 ? -dx : dx)) << 15) + (uint32_t)(dx + p.R);
    }
    const uint32_t rowbase = (uint32_t)(r * NN + (dy0 + p.R) * p.N);
    uint64_t acc[4][V];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int v = 0; v < V; v++) acc[k][v] = 0;
    auto span = [&](int k) { return *(const lds_u64_a4*)(rp + k); };
    uint64_t n0 = span(0), n1 = span(1), n2 = span(2), n3 = span(3);
#pragma clang loop unroll(full)
    for (int rho = 0; rho < V + 15; rho++) {
      uint64_t p0 = n0, p1 = n1, p2 = n2, p3 = n3;
      // the wait for this row's spans lands here, a row of qsads after
      // their loads were issued; later rows' loads stay below it
      asm volatile("" : "+v"(p0), "+v"(p1), "+v"(p2), "+v"(p3), "+v"(rp) : : "memory");
      if (rho + 1 < V + 15) {
        rp += p.row_dw;
        n0 = span(0); n1 = span(1); n2 = span(2); n3 = span(3);
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint64_t pk = k == 0 ? p0 : (k == 1 ? p1 : (k == 2 ? p2 : p3));
#pragma unroll
        for (int v = 0; v < V; v++) {
          const int br = rho - v;  // block row against reference row rho of this tile
          if (br >= 0 && br < 16) {
            const int quad = (br >= 8 ? 2 : 0) + (k >> 1);
            acc[quad][v] = __builtin_amdgcn_qsad_pk_u16_u8(pk, c[4 * br + k], acc[quad][v]);
          }
        }
      }
      // candidate row rho-15 has seen all 16 block rows.  Candidates outside
      // the MB's window get the key offset 2^31 instead of their rate and
      // index: keys stay below 2^32 (cost < 2^17), and the window always
      // holds the zero vector, whose key is below 2^31, so they never win.
      if (rho >= 15) {
        const int v = rho - 15;
        const int dy = dy0 + v;
        const bool row_ok = dy >= yu && dy <= yd;
        const uint32_t offr = (__umul24((uint32_t)p.lambda, (uint32_t)(dy < 0 ? -dy : dy)) << 15) +
                              rowbase + (uint32_t)(v * p.N);
        uint32_t off[4];
#pragma unroll
        for (int j = 0; j < 4; j++) off[j] = (row_ok && colok[j]) ? offr + offc[j] : 0x80000000u;
        // the nine partitions' SADs, two candidates (columns 2h, 2h + 1) per
        // dword: the quadrants as accumulated, the sums lane-wise
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const uint32_t s0 = (uint32_t)(acc[0][v] >> (32 * h)), s1 = (uint32_t)(acc[1][v] >> (32 * h));
          const uint32_t s2 = (uint32_t)(acc[2][v] >> (32 * h)), s3 = (uint32_t)(acc[3][v] >> (32 * h));
          const uint32_t top = add_u16x2(s0, s1), bot = add_u16x2(s2, s3);
          const uint32_t part[9] = {add_u16x2(top, bot), top, bot, add_u16x2(s0, s2), add_u16x2(s1, s3),
                                    s0, s1, s2, s3};
#pragma unroll
          for (int i = 0; i < 9; i++)
            best[i] = min3_u32(best[i], key_lo16(part[i], k15, off[2 * h]), key_hi16(part[i], k15, off[2 * h + 1]));
        }
      }
    }
  }
  // workgroup minimum of each partition's key
#pragma unroll
  for (int i = 0; i < 9; i++) {
    uint32_t k = best[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) k = min(k, (uint32_t)__shfl_xor((int)k, o, 64));
    if (lane == 0) red[wave][i] = k;
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t key[9];
    uint32_t J[9];
#pragma unroll
    for (int i = 0; i < 9; i++) {
      key[i] = min(min(red[0][i], red[1][i]), min(red[2][i], red[3][i]));
      J[i] = key[i] >> 15;
    }
    const uint32_t mj[4] = {J[0], J[1] + J[2], J[3] + J[4], J[5] + J[6] + J[7] + J[8]};
    int mode = 0;
#pragma unroll
    for (int k = 1; k < 4; k++)
      if (mj[k] < mj[mode]) mode = k;
    p.mode[mb] = (uint8_t)mode;
    p.cost[mb] = (int32_t)mj[mode];
    const int bw = p.W >> 3;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t idx = key[part_of_quadrant(mode, k)] & 0x7FFFu;
      const int r = (int)idx / NN, rem = (int)idx - r * NN;
      const int dy = rem / p.N - p.R, dx = rem - (rem / p.N) * p.N - p.R;
      const long b = (long)(2 * mrow + (k >> 1)) * bw + 2 * mcol + (k & 1);
      p.mv8[2 * b] = (int16_t)(4 * dx);
      p.mv8[2 * b + 1] = (int16_t)(4 * dy);
      p.ref8[b] = (uint8_t)r;
    }
  }
}

// ------------------------------------------------------- one MB per wave
// For ranges R = L0 or L0 - 1 with L0 = 8 or 16 (the window's half-width is
// R, clamped as GetMotionRange clamps it) one wave searches one MB over all
// references in turn.  The MB's block lives in SGPRs (every lane uses the
// same 64 dwords; qsad takes its block operand from an SGPR); each
// reference's window is staged with global_load_lds into the wave's own LDS
// and waited for with s_waitcnt, with no barrier.  Lanes tile the candidates
// from (-L0, -R) as QW column quads x BB row bands of VV rows, enough to
// cover the 2R + 1 columns and rows (lanes past QW x BB idle): 9 x 7 x 5 at
// R = 16, 8 x 8 x 4 at 15, 5 x 9 x 2 at 8, 4 x 16 x 1 at 7; candidates
// outside the MB's window get the key offset 2^31 and never win.  Per
// (reference row, block row) a lane issues 4 qsads into the accumulators of
// the quadrant that block row and column half belong to, and every finished
// candidate row folds into the lane's nine partition keys exactly as the
// workgroup kernel's; the wave's minima then decide the mode.
template <int L0, int QW, int BB, int VV>
struct VbsWaveGeom {
  static constexpr int SHIFT = ((16 - L0 % 16) % 16) / 4;  // dwords from the 16-byte boundary to column -L0
  static constexpr int ROW_DW = ((QW + 4 + SHIFT) + 3) / 4 * 4;
  static constexpr int ROWS = BB * VV + 15;
  static constexpr int WIN_DW = ROWS * ROW_DW;
  static_assert(QW * BB <= 64, "lane tiling");
};

template <int L0, int QW, int BB, int VV>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5)))
void vbs_wave_kernel(const VbsParams p) {
  using G = VbsWaveGeom<L0, QW, BB, VV>;
  __shared__ __attribute__((aligned(16))) uint32_t win_raw[G::WIN_DW];
  lds_u32* const win = (lds_u32*)win_raw;
  const int lane = threadIdx.x;
  int mb;
  {
    const int x = (int)blockIdx.x & 7, k = (int)blockIdx.x >> 3;
    const int q = p.nmb >> 3, r = p.nmb & 7;
    mb = x * q + min(x, r) + k;
  }
  const int mcol = mb % p.mbw, mrow = mb / p.mbw;
  const int n = 16 * mcol, m = 16 * mrow;
  const int R = p.R, N = p.N, NN = N * N;
  int xl, xr, yu, yd;
  motion_window(n, m, p.W, p.H, R, &xl, &xr, &yu, &yd);
  uint32_t best[9];
#pragma unroll
  for (int i = 0; i < 9; i++) best[i] = ~0u;

#pragma unroll 1
  for (int r = 0; r < p.nrefs; r++) {
    // Nothing but the keys lives across references: the block is loaded
    // again (uniform: scalar loads into SGPRs) and the lane index is made
    // opaque, so no per-lane value is hoisted out of the loop and spilled.
    int ln = lane;
    asm volatile("" : "+v"(ln));
    __builtin_assume(ln >= 0 && ln < 64);
    uint32_t c[64];
    {
      typedef __attribute__((address_space(4))) const uint32_t const_u32;
      const const_u32* cb = (const const_u32*)(p.cur + (long)m * p.W + n);
      const int sdw = p.W >> 2;
#pragma unroll
      for (int i = 0; i < 16; i++)
#pragma unroll
        for (int k = 0; k < 4; k++) c[4 * i + k] = cb[i * sdw + k];
    }
    const int q = ln % QW, b = ln / QW;
    const bool active = b < BB;
    const int dx0 = -L0 + 4 * q, dy0 = -R + (active ? b : 0) * VV;
    const uint32_t k15 = 1u << 15;
    uint32_t offc[4];
    bool colok[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int dx = dx0 + j;
      colok[j] = active && dx >= xl && dx <= xr;
      offc[j] = (__umul24((uint32_t)p.lambda, (uint32_t)(dx < 0 ? -dx : dx)) << 15) + (uint32_t)(dx + R);
    }
    const uint8_t* rb = r == 0 ? p.ref[0] : (r == 1 ? p.ref[1] : (r == 2 ? p.ref[2] : p.ref[3]));
    // stage rows m - R .. m - R + ROWS - 1 from the 16-byte boundary at or
    // left of column n - L0
    {
      // chunk base + ln: row ch / kCpr, 16-byte column ch % kCpr, at the
      // 32-bit offset row * stride + 16 (ch - kCpr row) from the uniform
      // origin (strides below 2^23)
      constexpr int kCpr = G::ROW_DW / 4, kChunks = G::ROWS * kCpr;
      const uint8_t* wsrc = rb + (long)(m - R) * p.ref_stride + n - L0 - 4 * G::SHIFT;
#pragma unroll
      for (int base = 0; base < kChunks; base += 64) {
        const uint32_t ch = (uint32_t)(base + ln);
        const uint32_t off = (uint32_t)__mul24((int)(ch / kCpr), p.ref_stride - 16 * kCpr) + (ch << 4);
        if (base + 64 <= kChunks || ch < (uint32_t)kChunks)
          __builtin_amdgcn_global_load_lds((const void*)(wsrc + off),
                                           (__attribute__((address_space(3))) void*)(win + 4 * base), 16, 0, 0);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t rowbase = __umul24((uint32_t)r, (uint32_t)NN) + __umul24((uint32_t)(dy0 + R), (uint32_t)N);
    uint64_t acc[4][VV];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int v = 0; v < VV; v++) acc[k][v] = 0;
    // idle lanes walk band 0's rows (in bounds) and fold nothing
    // P rows at a time are read at constant offsets from one base (the
    // offsets of ds_read2_b32 reach 255 dwords)
    constexpr int P = 251 / G::ROW_DW;
    const lds_u32* rp = win + G::SHIFT + (active ? b : 0) * VV * G::ROW_DW + q;
    asm volatile("" : "+v"(rp));
    auto span = [&](int rho, int k) { return *(const lds_u64_a4*)(rp + (rho % P) * G::ROW_DW + k); };
    uint64_t n0 = span(0, 0), n1 = span(0, 1), n2 = span(0, 2), n3 = span(0, 3);
#pragma clang loop unroll(full)
    for (int rho = 0; rho < VV + 15; rho++) {
      uint64_t p0 = n0, p1 = n1, p2 = n2, p3 = n3;
      asm volatile("" : "+v"(p0), "+v"(p1), "+v"(p2), "+v"(p3) : : "memory");
      if (rho + 1 < VV + 15) {
        if ((rho + 1) % P == 0) {
          rp += P * G::ROW_DW;
          asm volatile("" : "+v"(rp));
        }
        const int nr = rho + 1;
        n0 = span(nr, 0); n1 = span(nr, 1); n2 = span(nr, 2); n3 = span(nr, 3);
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint64_t pk = k == 0 ? p0 : (k == 1 ? p1 : (k == 2 ? p2 : p3));
#pragma unroll
        for (int v = 0; v < VV; v++) {
          const int br = rho - v;
          if (br >= 0 && br < 16) {
            const int quad = (br >= 8 ? 2 : 0) + (k >> 1);
            acc[quad][v] = __builtin_amdgcn_qsad_pk_u16_u8(pk, c[4 * br + k], acc[quad][v]);
          }
        }
      }
      if (rho >= 15) {
        const int v = rho - 15;
        const int dy = dy0 + v;
        const bool row_ok = dy >= yu && dy <= yd;
        const uint32_t offr = (__umul24((uint32_t)p.lambda, (uint32_t)(dy < 0 ? -dy : dy)) << 15) + rowbase +
                              __umul24((uint32_t)v, (uint32_t)N);
        uint32_t off[4];
#pragma unroll
        for (int j = 0; j < 4; j++) off[j] = (row_ok && colok[j]) ? offr + offc[j] : 0x80000000u;
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const uint32_t s0 = (uint32_t)(acc[0][v] >> (32 * h)), s1 = (uint32_t)(acc[1][v] >> (32 * h));
          const uint32_t s2 = (uint32_t)(acc[2][v] >> (32 * h)), s3 = (uint32_t)(acc[3][v] >> (32 * h));
          const uint32_t top = add_u16x2(s0, s1), bot = add_u16x2(s2, s3);
          const uint32_t part[9] = {add_u16x2(top, bot), top, bot, add_u16x2(s0, s2), add_u16x2(s1, s3),
                                    s0, s1, s2, s3};
#pragma unroll
          for (int i = 0; i < 9; i++)
            best[i] = min3_u32(best[i], key_lo16(part[i], k15, off[2 * h]), key_hi16(part[i], k15, off[2 * h + 1]));
        }
      }
    }
    // every lane's window reads are done before the next reference lands
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  // the wave's minimum of each partition's key (DPP, uniform)
#pragma unroll
  for (int i = 0; i < 9; i++) best[i] = wave_min_u32_uniform(best[i]);
  if (lane == 0) {
    uint32_t J[9];
#pragma unroll
    for (int i = 0; i < 9; i++) J[i] = best[i] >> 15;
    const uint32_t mj[4] = {J[0], J[1] + J[2], J[3] + J[4], J[5] + J[6] + J[7] + J[8]};
    int mode = 0;
#pragma unroll
    for (int k = 1; k < 4; k++)
      if (mj[k] < mj[mode]) mode = k;
    p.mode[mb] = (uint8_t)mode;
    p.cost[mb] = (int32_t)mj[mode];
    const int bw = p.W >> 3;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t idx = best[part_of_quadrant(mode, k)] & 0x7FFFu;
      const int rr = (int)idx / NN, rem = (int)idx - rr * NN;
      const int dy = rem / N - R, dx = rem - (rem / N) * N - R;
      const long b8 = (long)(2 * mrow + (k >> 1)) * bw + 2 * mcol + (k & 1);
      p.mv8[2 * b8] = (int16_t)(4 * dx);
      p.mv8[2 * b8 + 1] = (int16_t)(4 * dy);
      p.ref8[b8] = (uint8_t)rr;
    }
  }
}

// The lane tilings of the four instantiations (see above).
#define ICU_VBS_WAVE_R16 16, 9, 7, 5
#define ICU_VBS_WAVE_R15 16, 8, 8, 4
#define ICU_VBS_WAVE_R8 8, 5, 9, 2
#define ICU_VBS_WAVE_R7 8, 4, 16, 1

}  // namespace

int vbs_wave_class(int R) {
  for (int L0 : {8, 16})
    if (R == L0 || R == L0 - 1) return L0;
  return 0;
}

void vbs_plan(int R, int nrefs, VbsParams* p) {
  const int N = 2 * R + 1;
  p->R = R;
  p->N = N;
  p->nrefs = nrefs;
  p->dx0 = tile_dx0(R);
  p->Q = (R - p->dx0) / 4 + 1;
  // the column of n + dx0 within its 16-byte chunk (n is a multiple of 16)
  p->shift = ((p->dx0 % 16) + 16) % 16 / 4;
  // rows per lane tile: fewest task rounds of the workgroup, then the least
  // work per round (V qsad rows + the 15 extra rows every tile walks)
  int bestV = 4, bestCost = 1 << 30;
  for (int V = 4; V <= 8; V++) {
    const int Gr = (N + V - 1) / V;
    const int rounds = (nrefs * p->Q * Gr + kThreads - 1) / kThreads;
    const int cost = rounds * (V + 3);
    if (cost < bestCost) {
      bestCost = cost;
      bestV = V;
    }
  }
  p->V = bestV;
  p->Gr = (N + bestV - 1) / bestV;
  p->rows = p->Gr * bestV + 15;
  p->row_dw = ((p->shift + p->Q + 4) + 3) & ~3;
  p->wave = vbs_wave_class(R);
  // the kernel divides with udiv_small: at R <= 32 and 4 references the
  // chunk and task counts stay far inside u < 2^17, d <= 2048 (at most
  // 4 x 87 x 6 chunks, 4 x 17 x 17 tasks)
}

int vbs_lds_bytes(const VbsParams& p) {
  if (p.wave) {
    switch (p.R) {
      case 16: return VbsWaveGeom<ICU_VBS_WAVE_R16>::WIN_DW * 4;
      case 15: return VbsWaveGeom<ICU_VBS_WAVE_R15>::WIN_DW * 4;
      case 8: return VbsWaveGeom<ICU_VBS_WAVE_R8>::WIN_DW * 4;
      default: return VbsWaveGeom<ICU_VBS_WAVE_R7>::WIN_DW * 4;
    }
  }
  return p.nrefs * p.rows * p.row_dw * 4;
}
int vbs_tile_rows(const VbsParams& p) {
  if (!p.wave) return p.V;
  return p.R == 16 ? 5 : (p.R == 15 ? 4 : (p.R == 8 ? 2 : 1));
}

void launch_vbs_search(const VbsParams& p, hipStream_t st) {
  if (p.wave) {
    const dim3 g((unsigned)p.nmb), b(64);
    switch (p.R) {
      case 16: hipLaunchKernelGGL((vbs_wave_kernel<ICU_VBS_WAVE_R16>), g, b, 0, st, p); break;
      case 15: hipLaunchKernelGGL((vbs_wave_kernel<ICU_VBS_WAVE_R15>), g, b, 0, st, p); break;
      case 8: hipLaunchKernelGGL((vbs_wave_kernel<ICU_VBS_WAVE_R8>), g, b, 0, st, p); break;
      default: hipLaunchKernelGGL((vbs_wave_kernel<ICU_VBS_WAVE_R7>), g, b, 0, st, p); break;
    }
    return;
  }
  const size_t shmem = (size_t)vbs_lds_bytes(p);
  const dim3 grid((unsigned)p.nmb), block(kThreads);
  switch (p.V) {
    case 4: hipLaunchKernelGGL(vbs_search_kernel<4>, grid, block, shmem, st, p); break;
    case 5: hipLaunchKernelGGL(vbs_search_kernel<5>, grid, block, shmem, st, p); break;
    case 6: hipLaunchKernelGGL(vbs_search_kernel<6>, grid, block, shmem, st, p); break;
    case 7: hipLaunchKernelGGL(vbs_search_kernel<7>, grid, block, shmem, st, p); break;
    default: hipLaunchKernelGGL(vbs_search_kernel<8>, grid, block, shmem, st, p); break;
  }
}

}  // namespace icu
