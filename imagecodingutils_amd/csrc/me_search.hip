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

#include "icu_device.h"
#include "icu_internal.h"

namespace icu {

constexpr int kThreads = 256;
constexpr int kWinDwords = 8192;  // 32 KB staged reference window
constexpr int kMaxG = 16;         // MBs per task

// ---------------------------------------------------------------- padding
// Edge-replicated copy of an 8-bit plane into the padded plane
// (OverlayExtMem2Dv2::ExtendBoundary/FillBoundary,
// src/ImageUtils/OverlayExtMem2Dv2.cpp:120-217, as a coordinate clamp).
__global__ __launch_bounds__(256) void pad_plane_kernel(const uint8_t* __restrict__ src, int W,
                                                        int H, int src_stride,
                                                        uint8_t* __restrict__ dst_origin,
                                                        int dst_stride) {
  const int pw = W + 2 * kPad;
  const int x4 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;  // 4 bytes per thread
  const int y = blockIdx.y;
  if (x4 >= pw) return;
  const int yy = min(max(y - kPad, 0), H - 1);
  const uint8_t* row = src + (long)yy * src_stride;
  uint32_t w = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int xx = min(max(x4 + i - kPad, 0), W - 1);
    w |= (uint32_t)row[xx] << (8 * i);
  }
  *(uint32_t*)(dst_origin - (long)kPad * dst_stride - kPad + (long)y * dst_stride + x4) = w;
}

// ------------------------------------------------------------- tile bodies
// An opaque zero that depends on `dep`: adding it to an address keeps the
// load from being issued before `dep` is computed.
__device__ __forceinline__ uint32_t opaque_zero(uint32_t dep) {
  uint32_t z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z) : "v"(dep));
  return z;
}

// Rows are software-pipelined one ahead: the LDS loads of row rho+1 are
// issued before the qsads of row rho.  Their address carries an opaque
// dependency on row rho-1's sums, so the scheduler cannot hoist every row's
// loads to the top of the tile (which would exhaust the VGPRs).
template <int V, typename Done>
__device__ __forceinline__ void sad_tile(const uint32_t* wp, int row_dw,
                                         const uint32_t (&c)[64], Done&& done) {
  uint64_t acc[V];
#pragma unroll
  for (int v = 0; v < V; v++) acc[v] = 0;
  uint32_t n0 = wp[0], n1 = wp[1], n2 = wp[2], n3 = wp[3], n4 = wp[4];
  uint32_t dep = 0;
#pragma unroll
  for (int rho = 0; rho < V + 15; rho++) {
    const uint32_t d0 = n0, d1 = n1, d2 = n2, d3 = n3, d4 = n4;
    if (rho + 1 < V + 15) {
      const uint32_t* rp = wp + (rho + 1) * row_dw + opaque_zero(dep);
      n0 = rp[0]; n1 = rp[1]; n2 = rp[2]; n3 = rp[3]; n4 = rp[4];
    }
    const uint64_t p0 = ((uint64_t)d1 << 32) | d0;
    const uint64_t p1 = ((uint64_t)d2 << 32) | d1;
    const uint64_t p2 = ((uint64_t)d3 << 32) | d2;
    const uint64_t p3 = ((uint64_t)d4 << 32) | d3;
#pragma unroll
    for (int v = 0; v < V; v++) {
      const int r = rho - v;
      if (r >= 0 && r < 16) {
        acc[v] = __builtin_amdgcn_qsad_pk_u16_u8(p0, c[4 * r + 0], acc[v]);
        acc[v] = __builtin_amdgcn_qsad_pk_u16_u8(p1, c[4 * r + 1], acc[v]);
        acc[v] = __builtin_amdgcn_qsad_pk_u16_u8(p2, c[4 * r + 2], acc[v]);
        acc[v] = __builtin_amdgcn_qsad_pk_u16_u8(p3, c[4 * r + 3], acc[v]);
      }
    }
    dep = (uint32_t)acc[rho < V ? rho : V - 1];
    // accumulator rho-15 has seen all 16 block rows: retire it now so its
    // registers free up while the remaining rows run
    if (rho >= 15) done(rho - 15, (uint32_t)acc[rho - 15], (uint32_t)(acc[rho - 15] >> 32));
  }
}

// SSD tile: acc[v][k] = sum over rows of (sum r^2 - 2 sum c*r) for column
// offset k, in wrapping 32-bit arithmetic; the caller adds sum c^2.
template <int V, typename Done>
__device__ __forceinline__ void ssd_tile(const uint32_t* wp, int row_dw,
                                         const uint32_t (&c)[64], Done&& done) {
  uint32_t acc[V][4];
#pragma unroll
  for (int v = 0; v < V; v++)
#pragma unroll
    for (int k = 0; k < 4; k++) acc[v][k] = 0;
  uint32_t nx[5];
#pragma unroll
  for (int i = 0; i < 5; i++) nx[i] = wp[i];
  uint32_t dep = 0;
#pragma unroll
  for (int rho = 0; rho < V + 15; rho++) {
    uint32_t d[5];
#pragma unroll
    for (int i = 0; i < 5; i++) d[i] = nx[i];
    if (rho + 1 < V + 15) {
      const uint32_t* rp = wp + (rho + 1) * row_dw + opaque_zero(dep);
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
    dep = acc[rho < V ? rho : V - 1][0];
    if (rho >= 15) done(rho - 15, acc[rho - 15]);
  }
}

// ------------------------------------------------------------- the kernel
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

template <int METRIC, int V>
__device__ __forceinline__ void run_tile(const StreamDesc& s, const SearchTask& t, int u,
                                         const uint32_t* win, int row_dw, const uint32_t* curs,
                                         const uint32_t* csq, uint64_t* best, int* zero,
                                         uint32_t* __restrict__ table) {
  const int U = s.Q * s.B;
  const int g = u / U, rem = u - g * U;
  const int b = rem / s.Q, q = rem - b * s.Q;
  uint32_t c[64];
  const uint4* cp = (const uint4*)(curs + g * 64);
#pragma unroll
  for (int i = 0; i < 16; i++) {
    uint4 x = cp[i];
    c[4 * i + 0] = x.x;
    c[4 * i + 1] = x.y;
    c[4 * i + 2] = x.z;
    c[4 * i + 3] = x.w;
  }
  const uint32_t* wp = win + (b * V) * row_dw + 4 * g + q;
  const int mb = t.mb + g;
  const int col = mb % s.mbw, row = mb / s.mbw;
  const int n = col * 16, m = row * 16;
  int xl, xr, yu, yd;
  motion_window(n, m, s.W, s.H, s.L, &xl, &xr, &yu, &yd);
  const int dx0 = tile_dx0(s.L) + 4 * q;
  const int dy0 = -s.L + b * V;
  uint64_t kbest = ~0ull;
  const long long tdim = s.table_dim;
  uint32_t* trow = table ? table + s.table_base + (long long)mb * tdim * tdim : nullptr;
  if (METRIC == kMetricSAD) {
    sad_tile<V>(wp, row_dw, c, [&](int v, uint32_t lo, uint32_t hi) {
      visit(dx0 + 0, dy0 + v, lo & 0xFFFF, xl, xr, yu, yd, s.L, trow, tdim, &zero[g], kbest);
      visit(dx0 + 1, dy0 + v, lo >> 16, xl, xr, yu, yd, s.L, trow, tdim, &zero[g], kbest);
      visit(dx0 + 2, dy0 + v, hi & 0xFFFF, xl, xr, yu, yd, s.L, trow, tdim, &zero[g], kbest);
      visit(dx0 + 3, dy0 + v, hi >> 16, xl, xr, yu, yd, s.L, trow, tdim, &zero[g], kbest);
    });
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

template <int METRIC>
__global__ __launch_bounds__(kThreads) void me_search_kernel(const StreamDesc* __restrict__ streams,
                                                             const SearchTask* __restrict__ tasks,
                                                             int ntasks, IntResult* __restrict__ out,
                                                             uint32_t* __restrict__ table) {
  __shared__ uint32_t win[kWinDwords];
  __shared__ __attribute__((aligned(16))) uint32_t curs[kMaxG * 64];
  __shared__ uint32_t csq[kMaxG];
  __shared__ uint64_t best[kMaxG];
  __shared__ int zero[kMaxG];
  const int tid = threadIdx.x;

  for (int ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
    const SearchTask t = tasks[ti];
    const StreamDesc s = streams[t.stream];
    const int col0 = t.mb % s.mbw, row = t.mb / s.mbw;
    const int rows = s.B * s.V + 15;
    const int row_dw = 4 * (t.count - 1) + s.Q + 4;
    const uint8_t* wsrc = s.ref + (long)(16 * row - s.L) * s.ref_stride + 16 * col0 + tile_dx0(s.L);
    for (int i = tid; i < rows * row_dw; i += kThreads) {
      const int r = i / row_dw, cdw = i - r * row_dw;
      win[i] = *(const uint32_t*)(wsrc + (long)r * s.ref_stride + 4 * cdw);
    }
    for (int i = tid; i < t.count * 64; i += kThreads) {
      const int g = i >> 6, r = (i >> 2) & 15, k = i & 3;
      curs[i] = *(const uint32_t*)(s.cur + (long)(16 * row + r) * s.cur_stride + 16 * (col0 + g) + 4 * k);
    }
    if (tid < t.count) {
      best[tid] = ~0ull;
      zero[tid] = -1;
    }
    __syncthreads();
    if (METRIC == kMetricSSD && tid < t.count) {
      uint32_t a = 0;
      for (int i = 0; i < 64; i++) a = __builtin_amdgcn_udot4(curs[tid * 64 + i], curs[tid * 64 + i], a, false);
      csq[tid] = a;
    }
    __syncthreads();
    const int units = t.count * s.Q * s.B;
    for (int u = tid; u < units; u += kThreads) {
      if (METRIC == kMetricSSD) {
        switch (s.V) {
          case 8: run_tile<METRIC, 8>(s, t, u, win, row_dw, curs, csq, best, zero, table); break;
          case 11: run_tile<METRIC, 11>(s, t, u, win, row_dw, curs, csq, best, zero, table); break;
          case 13: run_tile<METRIC, 13>(s, t, u, win, row_dw, curs, csq, best, zero, table); break;
          default: run_tile<METRIC, 16>(s, t, u, win, row_dw, curs, csq, best, zero, table); break;
        }
      } else {
        switch (s.V) {
          case 8: run_tile<METRIC, 8>(s, t, u, win, row_dw, curs, csq, best, zero, table); break;
          case 11: run_tile<METRIC, 11>(s, t, u, win, row_dw, curs, csq, best, zero, table); break;
          case 13: run_tile<METRIC, 13>(s, t, u, win, row_dw, curs, csq, best, zero, table); break;
          default: run_tile<METRIC, 16>(s, t, u, win, row_dw, curs, csq, best, zero, table); break;
        }
      }
    }
    __syncthreads();
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
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------- host side
void launch_pad_plane(const uint8_t* src, int W, int H, int src_stride, uint8_t* dst_origin,
                      int dst_stride, hipStream_t st) {
  const int pw = W + 2 * kPad;
  dim3 grid((pw / 4 + 255) / 256, H + 2 * kPad);
  hipLaunchKernelGGL(pad_plane_kernel, grid, dim3(256), 0, st, src, W, H, src_stride, dst_origin,
                     dst_stride);
}

void launch_search(const StreamDesc* d_streams, const SearchTask* d_tasks, int ntasks, int grid,
                   int metric, IntResult* d_out, uint32_t* d_table, hipStream_t st) {
  if (grid <= 0 || grid > ntasks) grid = ntasks;
  if (metric == kMetricSAD)
    hipLaunchKernelGGL(me_search_kernel<kMetricSAD>, dim3(grid), dim3(kThreads), 0, st, d_streams,
                       d_tasks, ntasks, d_out, d_table);
  else
    hipLaunchKernelGGL(me_search_kernel<kMetricSSD>, dim3(grid), dim3(kThreads), 0, st, d_streams,
                       d_tasks, ntasks, d_out, d_table);
}

int search_max_window_dwords() { return kWinDwords; }
int search_max_group() { return kMaxG; }

}  // namespace icu
