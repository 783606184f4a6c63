// The reference area of one MB's search, staged in LDS, for the fast
// estimators (UMHS, Cross, FHS).
//
// Their per-MB walks read blocks at data-dependent positions, one stage
// after another; read from L2, every stage waits a round trip.  All but the
// starting points (the predictor and the previous frame's vector, which can
// lie anywhere within the picture's 16-pel border) lie inside the MB's
// search window, so the window's rows [m + yu - 3, m + yd + 19) and columns
// [n + xl - 3, n + xr + 19) (the 3-pel margins cover the sub-pel filters and
// 4-byte reads past a block's right edge) are loaded straight into LDS with
// global_load_lds at the start of the MB, before the MB waits for its
// neighbours, and every read inside them is an LDS read.  Reads outside go
// to the padded plane as before.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "icu_decide.h"
#include "icu_device.h"

namespace icu {
namespace fw {

// The largest window: L = 63 (motion range 64), rows 16 + 2 (63 + 3) = 148;
// columns from up to 15 pels left of the left margin to the right margin,
// 126 + 16 + 6 + 15 = 163, in whole 16-byte chunks = 176.  A 4-byte read
// can reach 4 bytes past the last row, hence the slack.
constexpr int kMaxCols = 176;
constexpr int kMaxRows = 148;
constexpr int kBytes = kMaxCols * kMaxRows + 16;

typedef __attribute__((address_space(3))) uint8_t lds_u8;
typedef __attribute__((address_space(3))) uint32_t lds_u32w;

struct Win {
  int x0, y0;      // padded-plane pel at LDS byte 0 (x0 a multiple of 16)
  int cols, rows;  // bytes per row (a multiple of 16), rows
  // a block of w x h pels at (x, y), with margin pels on every side, is staged
  __device__ __forceinline__ bool has(int x, int y, int w, int h, int margin) const {
    return x - margin >= x0 && y - margin >= y0 && x + w + margin <= x0 + cols && y + h + margin <= y0 + rows;
  }
};

__device__ __forceinline__ Win shape(int n, int m, int W, int H, int L) {
  int xl, xr, yu, yd;
  motion_window(n, m, W, H, L, &xl, &xr, &yu, &yd);
  Win w;
  w.x0 = (n + xl - 3) & ~15;  // floor to a 16-byte boundary (the padded plane's chunks)
  const int x1 = n + xr + 16 + 3;
  w.cols = ((x1 - w.x0) + 15) & ~15;
  w.y0 = m + yu - 3;
  w.rows = (yd - yu) + 16 + 6;
  return w;
}

// Issue the window's loads (16 bytes per lane per load); wait() before use.
__device__ __forceinline__ void issue(const DecideDesc& d, const Win& w, uint8_t* lds, int lane) {
  const int cpr = w.cols >> 4, total = cpr * w.rows;
  const uint8_t* src0 = d.ref + (long)w.y0 * d.ref_stride + w.x0;
  for (int base = 0; base < total; base += 64) {
    const int ch = base + lane;
    const int r = ch / cpr, c = ch - r * cpr;
    if (ch < total)
      __builtin_amdgcn_global_load_lds((const void*)(src0 + (long)r * d.ref_stride + 16 * c),
                                       (__attribute__((address_space(3))) void*)(lds + 16 * base), 16, 0, 0);
  }
}

__device__ __forceinline__ void wait() {
  __builtin_amdgcn_s_waitcnt(0);
  dec::wave_sync();
}

__device__ __forceinline__ int byte_at(const uint8_t* lds, const Win& w, int x, int y) {
  return lds[(y - w.y0) * w.cols + (x - w.x0)];
}

// This lane's 4 pels (block row lane / 4, columns 4 (lane % 4) .. +3) of the
// 16 x 16 block at (x, y), as one dword: from LDS when staged, else from
// the padded plane.
__device__ __forceinline__ uint32_t row4(const DecideDesc& d, const uint8_t* lds, const Win& w, bool staged, int x,
                                         int y, int lane) {
  const int xs = x + 4 * (lane & 3), yy = y + (lane >> 2), sh = xs & 3;
  const uint32_t* p;
  if (staged) p = (const uint32_t*)(lds + (yy - w.y0) * w.cols + (xs - sh - w.x0));
  else p = (const uint32_t*)(d.ref + (long)yy * d.ref_stride + (xs - sh));
  return __builtin_amdgcn_alignbyte(p[1], p[0], sh);
}

// The 22 x 22 patch whose (3, 3) is pel (x, y), into P, from LDS when
// staged (the caller checks w.has(x, y, 16, 16, 3)), else from the plane.
__device__ __forceinline__ void patch(const DecideDesc& d, const uint8_t* lds, const Win& w, bool staged, int x, int y,
                                      uint8_t* P, int lane) {
  if (!staged) {
    dec::load_patch(d, x, y, P, lane);
    return;
  }
  for (int i = lane; i < kPatch * kPatch; i += 64) {
    const int r = i / kPatch, c = i - r * kPatch;
    P[i] = lds[(y - 3 + r - w.y0) * w.cols + (x - 3 + c - w.x0)];
  }
}

}  // namespace fw
}  // namespace icu
