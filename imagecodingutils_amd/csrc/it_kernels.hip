// H.264 4x4 integer transform + quantisation, dequantisation + inverse
// transform, and the fused per-MB residual coder (luma motion compensation ->
// residual -> forward IT+Q -> levels -> dequant + inverse IT -> clipped
// reconstruction).  Each 4x4 block lives in registers of one lane.
//
// Arithmetic restated from:
//   src/CodecUtils/FastForward4x4ITImpl1.cpp:57-212,269-278 (forward; the
//     unrolled Impl2 and the MB-wide On16x16ITImpl1 give identical output)
//   src/CodecUtils/FastInverse4x4ITImpl1.cpp:60-236 (inverse, rows first)
//   src/CodecUtils/FastInverse4x4On16x16ITImpl1.cpp:120-360 (cols first)
//   src/ImageUtils/OverlayMem2Dv2.cpp:1195-1220,1343-1368 (Sub16x16,
//     Add16x16WithClip255)
//   src/CodecUtils/MotionCompensatorH264ImplStd.cpp:410-437 (luma compensate)
// Every (short) store of the reference is an int16 truncation here at the
// same point.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "icu_device.h"
#include "icu_internal.h"
#include "icu_qpel.h"
#include "icu_tx.h"

namespace icu {

namespace {

using namespace tx;

__global__ __launch_bounds__(256) void fwd_kernel(const int16_t* __restrict__ in, int16_t* __restrict__ out,
                                                  long n, int qp, int intra, int mode, int layout) {
  const long blk = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (blk >= n) return;
  int16_t b[16];
#pragma unroll
  for (int p = 0; p < 16; p++) b[p] = in[elem_off(blk, p, layout)];
  fwd4x4(b, qp, intra, mode);
#pragma unroll
  for (int p = 0; p < 16; p++) out[elem_off(blk, p, layout)] = b[p];
}

__global__ __launch_bounds__(256) void inv_kernel(const int16_t* __restrict__ in, int16_t* __restrict__ out,
                                                  long n, int qp, int mode, int layout, int cols_first,
                                                  const int* __restrict__ w) {
  const long blk = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (blk >= n) return;
  int16_t b[16];
#pragma unroll
  for (int p = 0; p < 16; p++) b[p] = in[elem_off(blk, p, layout)];
  int wl[16];
  if (w) {
#pragma unroll
    for (int p = 0; p < 16; p++) wl[p] = w[p];
  }
  inv4x4(b, qp, mode, cols_first, w ? wl : nullptr);
#pragma unroll
  for (int p = 0; p < 16; p++) out[elem_off(blk, p, layout)] = b[p];
}

// Fused coder: one workgroup of 256 lanes codes 16 MBs, one lane per 4x4.
constexpr int kMbPerWg = 16;

// The 4x4 luma prediction of block (bi, bj) of the MB at (n, m) for the
// quarter-pel vector (mvx, mvy), filtered in registers from the 9 x 9
// reference samples it needs (rows and columns -2 .. +6 of the block's
// full-pel origin), read as 9 x 3 aligned dwords from the padded plane: no
// LDS patch and no barrier.  The MB's full-pel origin is split and clamped
// into the padding (where rows and columns past the picture repeat its edge),
// so a clamped origin predicts what the far vector would.  Each of the 16 fraction cases computes only
// the half samples it needs (b, h, j of H.264 8.4.2.2.1, with the raw
// vertical sums shared across j's columns); the MB's 16 lanes take the same
// case.  pr[4 r + c]: the prediction at block row r, column c.
__device__ __forceinline__ void predict_block(const uint8_t* __restrict__ ref, int ref_stride, int W, int H, int n,
                                              int m, int mvx, int mvy, int bi, int bj, int (&pr)[16]) {
  const int fx = mvx / 4, fy = mvy / 4;
  int qx = mvx % 4, qy = mvy % 4, fox = 0, foy = 0;
  if (qx < 0) { fox = -1; qx += 4; }
  if (qy < 0) { foy = -1; qy += 4; }
  const int ox = min(max(n + fx + fox, -kPad + 3), W + kPad - 25);
  const int oy = min(max(m + fy + foy, -kPad + 3), H + kPad - 19);
  const int x0 = ox + 4 * bj - 2, y0 = oy + 4 * bi - 2;
  const int xa = x0 & ~3, sh = x0 - xa;
  const uint32_t* base = (const uint32_t*)(ref + (long)y0 * ref_stride + xa);
  const int sdw = ref_stride >> 2;
  uint32_t d[9][3];
#pragma unroll
  for (int k = 0; k < 9; k++)
#pragma unroll
    for (int q = 0; q < 3; q++) d[k][q] = base[k * sdw + q];
  // p[k][c]: reference sample (x0 + c, y0 + k); the block's full pels are p[2 + r][2 + c]
  int p[9][9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const uint32_t a0 = __builtin_amdgcn_alignbyte(d[k][1], d[k][0], sh);
    const uint32_t a1 = __builtin_amdgcn_alignbyte(d[k][2], d[k][1], sh);
    const uint32_t a2 = d[k][2] >> (8 * sh);
#pragma unroll
    for (int c = 0; c < 9; c++) {
      const uint32_t a = c < 4 ? a0 : (c < 8 ? a1 : a2);
      p[k][c] = (int)((a >> (8 * (c & 3))) & 0xFFu);
    }
  }
  // half samples: b right of (r, c), h below it, j diagonal
  auto G = [&](int r, int c) { return p[2 + r][2 + c]; };
  auto B = [&](int r, int c) {
    const int* q = p[2 + r];
    return clip255((ICU_TAP6(q[c], q[c + 1], q[c + 2], q[c + 3], q[c + 4], q[c + 5]) + 16) >> 5);
  };
  auto Hh = [&](int r, int c) {
    return clip255((ICU_TAP6(p[r][2 + c], p[r + 1][2 + c], p[r + 2][2 + c], p[r + 3][2 + c], p[r + 4][2 + c],
                             p[r + 5][2 + c]) + 16) >> 5);
  };
  // j needs the raw vertical sums of rows r .. r + 5 at columns c .. c + 5
  int J[16];
  auto jgrid = [&]() {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int v[9];
#pragma unroll
      for (int cc = 0; cc < 9; cc++)
        v[cc] = ICU_TAP6(p[r][cc], p[r + 1][cc], p[r + 2][cc], p[r + 3][cc], p[r + 4][cc], p[r + 5][cc]);
#pragma unroll
      for (int c = 0; c < 4; c++)
        J[4 * r + c] = clip255((ICU_TAP6(v[c], v[c + 1], v[c + 2], v[c + 3], v[c + 4], v[c + 5]) + 512) >> 10);
    }
  };
#define ICU_PRED(expr)                  \
  _Pragma("unroll") for (int r = 0; r < 4; r++) \
    _Pragma("unroll") for (int c = 0; c < 4; c++) pr[4 * r + c] = (expr)
  switch (qx | (qy << 2)) {
    case 0: ICU_PRED(G(r, c)); break;
    case 2: ICU_PRED(B(r, c)); break;
    case 8: ICU_PRED(Hh(r, c)); break;
    case 10: jgrid(); ICU_PRED(J[4 * r + c]); break;
    case 1: ICU_PRED((G(r, c) + B(r, c) + 1) >> 1); break;
    case 3: ICU_PRED((G(r, c + 1) + B(r, c) + 1) >> 1); break;
    case 4: ICU_PRED((G(r, c) + Hh(r, c) + 1) >> 1); break;
    case 12: ICU_PRED((G(r + 1, c) + Hh(r, c) + 1) >> 1); break;
    case 6: jgrid(); ICU_PRED((J[4 * r + c] + B(r, c) + 1) >> 1); break;
    case 14: jgrid(); ICU_PRED((J[4 * r + c] + B(r + 1, c) + 1) >> 1); break;
    case 9: jgrid(); ICU_PRED((J[4 * r + c] + Hh(r, c) + 1) >> 1); break;
    case 11: jgrid(); ICU_PRED((J[4 * r + c] + Hh(r, c + 1) + 1) >> 1); break;
    case 5: ICU_PRED((B(r, c) + Hh(r, c) + 1) >> 1); break;
    case 7: ICU_PRED((B(r, c) + Hh(r, c + 1) + 1) >> 1); break;
    case 13: ICU_PRED((Hh(r, c) + B(r + 1, c) + 1) >> 1); break;
    default: ICU_PRED((Hh(r, c + 1) + B(r + 1, c) + 1) >> 1); break;
  }
#undef ICU_PRED
}

// The coders' levels and reconstructions are written once and read by a
// later launch or the host: nontemporal stores keep them from displacing
// the reference and input planes in the caches (a 1080p frame at three QPs
// writes 18.8 MB; the coder measured 10.8 us with ordinary stores, 9.5 us
// with these).
typedef unsigned int u32x2_nt __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void stream_store(uint2 v, int16_t* p) {
  __builtin_nontemporal_store((u32x2_nt){v.x, v.y}, (u32x2_nt*)p);
}

__global__ __launch_bounds__(256) void code_frame_kernel(const uint8_t* __restrict__ cur, int cur_stride,
                                                         const uint8_t* __restrict__ ref, int ref_stride,
                                                         int W, int H, const int16_t* __restrict__ mv,
                                                         const QpSet qs, int intra, int cols_first,
                                                         int16_t* __restrict__ pred_out) {
  // one lane per 4x4 block; the 16 lanes of an MB are consecutive
  const int mbw = W / 16, nmb = mbw * (H / 16);
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int mb = gid >> 4, blk = gid & 15;
  if (mb >= nmb) return;
  const int n = (mb % mbw) * 16, m = (mb / mbw) * 16;
  const int bi = blk >> 2, bj = blk & 3;
  uint32_t cw[4];
#pragma unroll
  for (int rr = 0; rr < 4; rr++)
    cw[rr] = *(const uint32_t*)(cur + (long)(m + 4 * bi + rr) * cur_stride + n + 4 * bj);
  int prd[16];
  predict_block(ref, ref_stride, W, H, n, m, mv[2 * mb], mv[2 * mb + 1], bi, bj, prd);
  int16_t b[16], pr[16];
#pragma unroll
  for (int p = 0; p < 16; p++) {
    pr[p] = (int16_t)prd[p];
    b[p] = (int16_t)((int)((cw[p >> 2] >> (8 * (p & 3))) & 0xFF) - pr[p]);
  }
  // 4 samples per row: 8-byte level / prediction stores, 4-byte recon stores
  // (W and n are multiples of 16)
  if (pred_out) {
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const long o = (long)(m + 4 * bi + rr) * W + n + 4 * bj;
      uint2 pv;
      pv.x = (uint32_t)(uint16_t)pr[4 * rr] | ((uint32_t)(uint16_t)pr[4 * rr + 1] << 16);
      pv.y = (uint32_t)(uint16_t)pr[4 * rr + 2] | ((uint32_t)(uint16_t)pr[4 * rr + 3] << 16);
      *(uint2*)(pred_out + o) = pv;
    }
  }
  // the residual is coded once per QP of the set
#pragma unroll 1
  for (int iq = 0; iq < qs.n; iq++) {
    const int qp = qs.qp[iq];
    int16_t t[16];
#pragma unroll
    for (int p = 0; p < 16; p++) t[p] = b[p];
    fwd4x4(t, qp, intra, 0);
    int16_t* levels = qs.levels[iq];
    uint8_t* recon = qs.recon[iq];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const long o = (long)(m + 4 * bi + rr) * W + n + 4 * bj;
      uint2 lv;
      lv.x = (uint32_t)(uint16_t)t[4 * rr] | ((uint32_t)(uint16_t)t[4 * rr + 1] << 16);
      lv.y = (uint32_t)(uint16_t)t[4 * rr + 2] | ((uint32_t)(uint16_t)t[4 * rr + 3] << 16);
      stream_store(lv, levels + o);
    }
    inv4x4(t, qp, 0, cols_first, nullptr);
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const long o = (long)(m + 4 * bi + rr) * W + n + 4 * bj;
      uint32_t rc = 0;
#pragma unroll
      for (int cc = 0; cc < 4; cc++) rc |= (uint32_t)clip255(pr[4 * rr + cc] + t[4 * rr + cc]) << (8 * cc);
      __builtin_nontemporal_store(rc, (uint32_t*)(recon + o));
    }
  }
}

}  // namespace

void launch_fwd4x4(const int16_t* in, int16_t* out, long n, int qp, int intra, int mode, int layout,
                   hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fwd_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, out, n, qp,
                     intra, mode, layout);
}

void launch_inv4x4(const int16_t* in, int16_t* out, long n, int qp, int mode, int layout,
                   int cols_first, const int* d_weights, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(inv_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, out, n, qp,
                     mode, layout, cols_first, d_weights);
}

namespace {
// Residual coding of a frame predicted per 8x8 block at an integer vector
// from one of several padded references (the multi-reference search's
// output): one lane per 4x4 block.  The prediction is a copy
// (MotionCompensatorH264ImplStd::Compensate at a full-pel vector,
// src/CodecUtils/MotionCompensatorH264ImplStd.cpp:410-437); the coding at
// each QP of the set is code_frame_kernel's.
__global__ __launch_bounds__(256) void code_vbs_kernel(const uint8_t* __restrict__ cur, const VbsRefs refs,
                                                       int W, int H, const int16_t* __restrict__ mv8,
                                                       const uint8_t* __restrict__ ref8, const QpSet qs,
                                                       int intra, int cols_first,
                                                       int16_t* __restrict__ pred_out) {
  const int bw4 = W >> 2;
  const long blk = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (blk >= (long)bw4 * (H >> 2)) return;
  const int by = (int)(blk / bw4), bx = (int)(blk - (long)by * bw4);
  const int x = 4 * bx, y = 4 * by;
  const long b8 = (long)(by >> 1) * (W >> 3) + (bx >> 1);
  const int ri = ref8[b8];
  const uint8_t* rb = ri == 0 ? refs.ref[0] : (ri == 1 ? refs.ref[1] : (ri == 2 ? refs.ref[2] : refs.ref[3]));
  // Full-pel source of the block (vectors floored to full pels), clamped
  // into the padding: beyond the picture the padded plane repeats its edge,
  // so a block clamped there reads what the far one would.  4 prediction
  // rows of 4 bytes at a byte offset within 2 aligned dwords (the padded
  // planes' pixel (0,0) is 16-byte aligned).
  const int xs = min(max(x + (mv8[2 * b8] >> 2), -kPad + 4), W + kPad - 8);
  const int ys = min(max(y + (mv8[2 * b8 + 1] >> 2), -kPad), H + kPad - 4);
  const int sh = xs & 3;
  uint32_t pw[4], cw[4];
#pragma unroll
  for (int rr = 0; rr < 4; rr++) {
    const uint32_t* a = (const uint32_t*)(rb + (long)(ys + rr) * refs.stride + (xs - sh));
    pw[rr] = __builtin_amdgcn_alignbyte(a[1], a[0], sh);
    cw[rr] = *(const uint32_t*)(cur + (long)(y + rr) * W + x);
  }
  int16_t b[16], pr[16];
#pragma unroll
  for (int p = 0; p < 16; p++) {
    pr[p] = (int16_t)((pw[p >> 2] >> (8 * (p & 3))) & 0xFF);
    b[p] = (int16_t)((int)((cw[p >> 2] >> (8 * (p & 3))) & 0xFF) - pr[p]);
  }
  if (pred_out) {
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      uint2 pv;
      pv.x = (uint32_t)(uint16_t)pr[4 * rr] | ((uint32_t)(uint16_t)pr[4 * rr + 1] << 16);
      pv.y = (uint32_t)(uint16_t)pr[4 * rr + 2] | ((uint32_t)(uint16_t)pr[4 * rr + 3] << 16);
      *(uint2*)(pred_out + (long)(y + rr) * W + x) = pv;
    }
  }
#pragma unroll 1
  for (int iq = 0; iq < qs.n; iq++) {
    const int qp = qs.qp[iq];
    int16_t t[16];
#pragma unroll
    for (int p = 0; p < 16; p++) t[p] = b[p];
    fwd4x4(t, qp, intra, 0);
    int16_t* levels = qs.levels[iq];
    uint8_t* recon = qs.recon[iq];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      uint2 lv;
      lv.x = (uint32_t)(uint16_t)t[4 * rr] | ((uint32_t)(uint16_t)t[4 * rr + 1] << 16);
      lv.y = (uint32_t)(uint16_t)t[4 * rr + 2] | ((uint32_t)(uint16_t)t[4 * rr + 3] << 16);
      stream_store(lv, levels + (long)(y + rr) * W + x);
    }
    inv4x4(t, qp, 0, cols_first, nullptr);
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      uint32_t rc = 0;
#pragma unroll
      for (int cc = 0; cc < 4; cc++) rc |= (uint32_t)clip255(pr[4 * rr + cc] + t[4 * rr + cc]) << (8 * cc);
      __builtin_nontemporal_store(rc, (uint32_t*)(recon + (long)(y + rr) * W + x));
    }
  }
}

// Whole-frame YUV 4:2:0 compensation, MotionCompensatorH264ImplStd::Compensate
// (src/CodecUtils/MotionCompensatorH264ImplStd.cpp:323-473): 16 MBs per
// workgroup.  Lane blk of an MB predicts its 4x4 luma block at the quarter-pel
// vector (QuarterRead from the staged patch) and 4 samples of row blk / 2 of
// each 8x8 chroma block at the same vector read as eighth pels
// (OverlayMem2Dv2::EighthRead, src/ImageUtils/OverlayMem2Dv2.cpp:1063-1114:
// C-truncated full / eighth split, negative fractions one pel left / above
// with 8 + f, bilinear ((8-x)(8-y)A + x(8-y)B + (8-x)yC + xyD + 32) >> 6,
// zero fractions copy).  Sources are the padded copies, so the outputs may
// be the reference planes themselves (the reference compensates in place
// from a copy).
__global__ __launch_bounds__(256) void mc_yuv_kernel(const uint8_t* __restrict__ ref_y, int ys,
                                                     const uint8_t* __restrict__ ref_u,
                                                     const uint8_t* __restrict__ ref_v, int cs, int W, int H,
                                                     const int16_t* __restrict__ mv, uint8_t* out_y,
                                                     uint8_t* out_u, uint8_t* out_v) {
  const int mbw = W / 16, nmb = mbw * (H / 16);
  const int lmb = threadIdx.x >> 4, blk = threadIdx.x & 15;
  const int mb = blockIdx.x * kMbPerWg + lmb;
  if (mb >= nmb) return;
  const int n = (mb % mbw) * 16, m = (mb / mbw) * 16;
  const int mvx = mv[2 * mb], mvy = mv[2 * mb + 1];
  uint32_t cu[4], cv[4];
  int cfx = 0, cfy = 0, csh = 0;
  {
    // chroma: rows oy + r, oy + r + 1, columns ox + c0 .. + 4
    const int Wc = W >> 1, Hc = H >> 1;
    cfx = mvx % 8;
    cfy = mvy % 8;
    int ox = (n >> 1) + mvx / 8, oy = (m >> 1) + mvy / 8;
    if (cfx < 0) { ox -= 1; cfx += 8; }
    if (cfy < 0) { oy -= 1; cfy += 8; }
    // clamped into the padding: beyond it every read would repeat the edge
    ox = min(max(ox, -kPad + 3), Wc + kPad - 12);
    oy = min(max(oy, -kPad), Hc + kPad - 9);
    const int r = blk >> 1, xs = ox + 4 * (blk & 1);
    csh = xs & 3;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const long o = (long)(oy + r + k) * cs + (xs - csh);
      cu[2 * k] = *(const uint32_t*)(ref_u + o);
      cu[2 * k + 1] = *(const uint32_t*)(ref_u + o + 4);
      cv[2 * k] = *(const uint32_t*)(ref_v + o);
      cv[2 * k + 1] = *(const uint32_t*)(ref_v + o + 4);
    }
  }
  // luma 4x4 block
  const int bi = blk >> 2, bj = blk & 3;
  int prd[16];
  predict_block(ref_y, ys, W, H, n, m, mvx, mvy, bi, bj, prd);
#pragma unroll
  for (int rr = 0; rr < 4; rr++) {
    uint32_t o = 0;
#pragma unroll
    for (int cc = 0; cc < 4; cc++) o |= (uint32_t)prd[4 * rr + cc] << (8 * cc);
    *(uint32_t*)(out_y + (long)(m + 4 * bi + rr) * W + n + 4 * bj) = o;
  }
  // chroma: 4 samples of one row per plane
  const long co = (long)((m >> 1) + (blk >> 1)) * (W >> 1) + (n >> 1) + 4 * (blk & 1);
  auto chroma = [&](const uint32_t (&w)[4]) {
    const uint64_t r0 = ((uint64_t)w[1] << 32 | w[0]) >> (8 * csh);
    const uint64_t r1 = ((uint64_t)w[3] << 32 | w[2]) >> (8 * csh);
    uint32_t o = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int A = (int)((r0 >> (8 * c)) & 255), B = (int)((r0 >> (8 * c + 8)) & 255);
      const int C = (int)((r1 >> (8 * c)) & 255), D = (int)((r1 >> (8 * c + 8)) & 255);
      const int v = (cfx | cfy) ? ((8 - cfx) * (8 - cfy) * A + cfx * (8 - cfy) * B + (8 - cfx) * cfy * C +
                                   cfx * cfy * D + 32) >> 6
                                : A;
      o |= (uint32_t)v << (8 * c);
    }
    return o;
  };
  *(uint32_t*)(out_u + co) = chroma(cu);
  *(uint32_t*)(out_v + co) = chroma(cv);
}

// Order-free checksum of a byte array: the wrapping 64-bit sum over i of
// (p[i] + 1) * (odd 32-bit hash of i), so it does not depend on which thread
// adds which part, and a changed or moved byte changes it.
__global__ __launch_bounds__(256) void checksum_kernel(const uint8_t* __restrict__ p, long n,
                                                       unsigned long long* __restrict__ out) {
  __shared__ unsigned long long part[4];
  unsigned long long s = 0;
  const long stride = (long)gridDim.x * blockDim.x;
  const long ndw = n >> 2;  // whole dwords (p is 4-byte aligned)
  const uint32_t* pw = (const uint32_t*)p;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < ndw; k += stride) {
    const uint32_t w = pw[k];
    const uint32_t i0 = (uint32_t)k * 4u;
#pragma unroll
    for (int j = 0; j < 4; j++)
      s += (unsigned long long)(((w >> (8 * j)) & 255u) + 1u) * (((i0 + j) * 0x9E3779B1u) | 1u);
  }
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const long i = (ndw << 2) + threadIdx.x;
    s += (unsigned long long)(p[i] + 1u) * (((uint32_t)i * 0x9E3779B1u) | 1u);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, part[0] + part[1] + part[2] + part[3]);
}
}  // namespace

void launch_checksum(const void* p, long bytes, unsigned long long* out, hipStream_t st) {
  // one atomic per workgroup: few workgroups, each thread over many dwords
  const long blocks = std::min<long>(128, (bytes / 4 + 255) / 256);
  if (bytes <= 0) return;
  hipLaunchKernelGGL(checksum_kernel, dim3((unsigned)std::max<long>(1, blocks)), dim3(256), 0, st,
                     (const uint8_t*)p, bytes, out);
}

void launch_mc_yuv(const uint8_t* ref_y, int ys, const uint8_t* ref_u, const uint8_t* ref_v, int cs, int W,
                   int H, const int16_t* mv, uint8_t* out_y, uint8_t* out_u, uint8_t* out_v, hipStream_t st) {
  const int nmb = (W / 16) * (H / 16);
  hipLaunchKernelGGL(mc_yuv_kernel, dim3((nmb + kMbPerWg - 1) / kMbPerWg), dim3(256), 0, st, ref_y, ys, ref_u,
                     ref_v, cs, W, H, mv, out_y, out_u, out_v);
}

void launch_code_vbs(const uint8_t* cur, const VbsRefs& refs, int W, int H, const int16_t* mv8,
                     const uint8_t* ref8, const QpSet& qs, int intra, int cols_first,
                     int16_t* pred_out, hipStream_t st) {
  const long nblk = (long)(W / 4) * (H / 4);
  hipLaunchKernelGGL(code_vbs_kernel, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, st, cur, refs, W,
                     H, mv8, ref8, qs, intra, cols_first, pred_out);
}

void launch_code_frame(const uint8_t* cur, int cur_stride, const uint8_t* ref_padded, int ref_stride,
                       int W, int H, const int16_t* mv, const QpSet& qs, int intra, int cols_first,
                       int16_t* pred_out, hipStream_t st) {
  const int nmb = (W / 16) * (H / 16);
  hipLaunchKernelGGL(code_frame_kernel, dim3((nmb + kMbPerWg - 1) / kMbPerWg), dim3(256), 0, st, cur,
                     cur_stride, ref_padded, ref_stride, W, H, mv, qs, intra, cols_first, pred_out);
}

}  // namespace icu
