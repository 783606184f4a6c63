// Shared device-side types for the motion-estimation and transform kernels.
// Host code fills these descriptors; kernels read them from device memory.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icu {

// Padding of the device reference plane.  Block origins searched lie in
// [-16, W] x [-16, H] (the reference's 16-pel border rule); the search tiles
// round the window up to whole lane tiles, and the sub-pel stages read up to
// 3 pels beyond a block, so 80 pels on every side keep every read in bounds
// for search ranges up to 64.
constexpr int kPad = 80;
constexpr int kMaxRange = 64;

// Window half-width rule.
enum { kWindowRef = 0, kWindowFull = 1 };
enum { kMetricSSD = 0, kMetricSAD = 1 };

// One luma frame pair being estimated.  All pointers are device pointers.
struct StreamDesc {
  const uint8_t* cur;   // W x H, row stride cur_stride
  const uint8_t* ref;   // padded plane; points at pixel (0,0), stride ref_stride
  int cur_stride, ref_stride;
  int W, H, mbw, mbh;
  int R;                // search range (full pel)
  int L;                // window half-width: R - 1 (reference rule) or R
  int V;                // candidate rows per lane (tile height), chosen per stream
  int B;                // ceil((2L+1) / V) row tiles
  int Q;                // ceil((2L+1) / 4) column quads
  int G;                // MBs per workgroup task
  int metric;
  int mb_base;          // first global MB slot of this stream in the outputs
  int table_base_valid; // 1 when a cost table is written for this stream
  long long table_base; // offset of this stream's first table entry
  int table_dim;        // 2L+1 (table rows and columns)
};

// A workgroup task: `count` consecutive MBs of one MB row of one stream.
struct SearchTask {
  int stream;
  int mb;     // first MB (raster index within the stream)
  int count;  // MBs in this task (1..G)
};

// Integer-stage result per MB, packed for a 64-bit atomic min:
//   bits 63..32  cost (SAD <= 65280, SSD <= 16,646,400)
//   bits 31..17  squared distance to the rounded predictor (zero here)
//   bits 16..0   reverse raster rank (later candidates rank lower; the zero
//                vector, the initial incumbent, ranks highest)
// The minimum key is the reference's decision with the running-best filter
// off: lower cost wins, then the candidate closer to the predictor, then the
// later candidate in raster order.
__host__ __device__ inline uint64_t pack_key(uint32_t cost, uint32_t dist2, uint32_t rrank) {
  return ((uint64_t)cost << 32) | ((uint64_t)(dist2 & 0x7FFF) << 17) | (rrank & 0x1FFFF);
}
__host__ __device__ inline uint32_t raster_rrank(int i, int j) {
  // i, j in [-64, 64]; rank increases in raster order (row i, then column j).
  return 0xFFFFu - (uint32_t)((i + 64) * 256 + (j + 64));
}
constexpr uint32_t kZeroRRank = 0x1FFFF;
__host__ __device__ inline uint32_t key_cost(uint64_t k) { return (uint32_t)(k >> 32); }
// Vector of a key: the zero vector for kZeroRRank, else the inverse of raster_rrank.
__host__ __device__ inline void key_vector(uint64_t k, int* dx, int* dy) {
  uint32_t rr = (uint32_t)(k & 0x1FFFF);
  if (rr == kZeroRRank) { *dx = 0; *dy = 0; return; }
  uint32_t rank = 0xFFFFu - rr;
  *dy = (int)(rank >> 8) - 64;
  *dx = (int)(rank & 255) - 64;
}

// The clamped search window of the MB at pel (n, m): the reference's
// GetMotionRange with offset 0 (half-width L, blocks may reach 16 pels
// outside the picture, never further).
__host__ __device__ inline void motion_window(int n, int m, int W, int H, int L, int* xl,
                                              int* xr, int* yu, int* yd) {
  *xl = (n - L >= -16) ? -L : -(n + 16);
  *xr = (n + L < W) ? L : W - n;
  *yu = (m - L >= -16) ? -L : -(m + 16);
  *yd = (m + L < H) ? L : H - m;
}

// The least 32-bit value over the wave, in every lane as a uniform value:
// row_ror 1, 2, 4, 8 leave each lane its row's minimum, row_bcast:15 and
// row_bcast:31 carry rows 0 and 2 into 1 and 3 and lane 31 into rows 2 and
// 3, so lane 63 holds the wave's (lanes outside a row mask read the
// identity, ~0, so each step is one v_min_u32 with DPP).
__device__ __forceinline__ uint32_t wave_min_u32_uniform(uint32_t v) {
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x121, 0xf, 0xf, false));  // row_ror:1
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x122, 0xf, 0xf, false));  // row_ror:2
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x124, 0xf, 0xf, false));  // row_ror:4
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x128, 0xf, 0xf, false));  // row_ror:8
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
// First candidate column covered by the lane tiles: -L rounded down to a
// multiple of 4 so every lane's 4-candidate quad starts on a dword.
__host__ __device__ inline int tile_dx0(int L) { return -4 * ((L + 3) / 4); }

// u / d for 0 <= u < 2^17 and 1 <= d <= 2048 without an integer division:
// (u + 1/2) / d lies at least 1/(2d) from an integer, and the f32 product
// with v_rcp_f32's reciprocal (within 1 ulp of 1/d), rounded once or twice,
// stays closer than that; checked exhaustively over the whole domain with the
// reciprocal at 1/d and at both neighbouring floats.
__device__ __forceinline__ int udiv_small(int u, int d) {
  return (int)(((float)u + 0.5f) * __builtin_amdgcn_rcpf((float)d));
}

// Lane tiles of the search kernel's argmin keys: tiles taller than 8 rows
// keep one key per half, rows [0, V1) and [V1, V).
__host__ __device__ constexpr int lane_key_split(int V) { return V > 8 ? (V + 1) / 2 : V; }
// Bits of a tile half's raster index (4 V1 candidates) in those keys.
__host__ __device__ constexpr int lane_key_bits(int V) { return 4 * lane_key_split(V) <= 32 ? 5 : 6; }

// True when every lane tile half of the window (half-width L, tile rows V)
// keeps its argmin codes ((dist2 - least dist2 of the half) << bits | raster
// part) below 2^16, so that (cost << 16) | code orders the half exactly.
inline bool lane_keys_fit(int L, int V) {
  const int ib = lane_key_bits(V), V1 = lane_key_split(V);
  const int N = 2 * L + 1, B = (N + V - 1) / V, Q = (L - tile_dx0(L)) / 4 + 1;
  auto span = [](int a, int len, int* lo2, int* hi2) {  // min / max of x^2 on [a, a+len)
    const int b = a + len - 1;
    const int m = a > 0 ? a : (b < 0 ? -b : 0);
    const int M = (a < 0 ? -a : a) > (b < 0 ? -b : b) ? (a < 0 ? -a : a) : (b < 0 ? -b : b);
    *lo2 = m * m;
    *hi2 = M * M;
  };
  for (int b = 0; b < B; b++)
    for (int h = 0; h < (V1 < V ? 2 : 1); h++)
      for (int q = 0; q < Q; q++) {
        int xl2, xh2, yl2, yh2;
        span(tile_dx0(L) + 4 * q, 4, &xl2, &xh2);
        span(-L + b * V + h * V1, h ? V - V1 : V1, &yl2, &yh2);
        const long code = ((long)(xh2 + yh2 - xl2 - yl2) << ib) + (1 << ib) - 1;
        if (code > 0xFFFF) return false;
      }
  return true;
}

// Per-MB integer-stage output of the search kernel.
struct IntResult {
  int32_t mx, my;   // full-pel vector
  int32_t cost;     // its cost
  int32_t zero;     // cost of the zero vector
};

}  // namespace icu
