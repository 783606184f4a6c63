// H.264 4x4 forward integer transform + quantisation and dequantisation +
// inverse transform of one block, as host/device code: the kernels keep a
// block in one lane's registers, and host callers of single blocks run the
// same arithmetic on the CPU.
//
// Arithmetic restated from:
//   src/CodecUtils/FastForward4x4ITImpl1.cpp:57-212,269-278 (forward; the
//     unrolled Impl2 and the MB-wide On16x16ITImpl1 give identical output)
//   src/CodecUtils/FastInverse4x4ITImpl1.cpp:60-236 (inverse, rows first)
//   src/CodecUtils/FastInverse4x4On16x16ITImpl1.cpp:120-360 (cols first)
// Every (short) store of the reference is an int16 truncation here at the
// same point.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ICU_HD __host__ __device__ __forceinline__

namespace icu {
namespace tx {

// Quantiser multipliers MF and dequantiser scales V per qp % 6 and position
// class (H.264 8.5.12)
constexpr int kMF[6][3] = {{13107, 5243, 8066}, {11916, 4660, 7490}, {10082, 4194, 6554},
                           {9362, 3647, 5825},  {8192, 3355, 5243},  {7282, 2893, 4559}};
constexpr int kV[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16},
                          {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};

ICU_HD int pos_class(int p) {
  // 0 = (even,even), 1 = (odd,odd), 2 = mixed
  const int r = p >> 2, c = p & 3;
  return ((r | c) & 1) == 0 ? 0 : (((r & c) & 1) ? 1 : 2);
}

ICU_HD int16_t quant(int x, int p, int qm, int f, int scale) {
  const int mf = kMF[qm][pos_class(p)];
  if (x < 0) return (int16_t)(-((((-x) * mf) + f) >> scale));
  return (int16_t)(((x * mf) + f) >> scale);
}

ICU_HD void fwd4x4(int16_t (&b)[16], int qp, int intra, int mode) {
  const int qm = qp % 6, qe = qp / 6, scale = 15 + qe;
  const int f = intra ? (1 << (15 + qe)) / 3 : (1 << (15 + qe)) / 6;
  if (mode == 2) {
#pragma unroll
    for (int p = 0; p < 16; p++) b[p] = quant(b[p], p, qm, f, scale);
    return;
  }
#pragma unroll
  for (int j = 0; j < 16; j += 4) {
    const int s0 = b[j] + b[j + 3], s3 = b[j] - b[j + 3];
    const int s1 = b[j + 1] + b[j + 2], s2 = b[j + 1] - b[j + 2];
    b[j] = (int16_t)(s0 + s1);
    b[j + 2] = (int16_t)(s0 - s1);
    b[j + 1] = (int16_t)(s2 + s3 * 2);
    b[j + 3] = (int16_t)(s3 - s2 * 2);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int s0 = b[j] + b[j + 12], s3 = b[j] - b[j + 12];
    const int s1 = b[j + 4] + b[j + 8], s2 = b[j + 4] - b[j + 8];
    const int x0 = s0 + s1, x2 = s0 - s1, x1 = s2 + s3 * 2, x3 = s3 - s2 * 2;
    if (mode == 0) {
      b[j] = quant(x0, j, qm, f, scale);
      b[j + 8] = quant(x2, j + 8, qm, f, scale);
      b[j + 4] = quant(x1, j + 4, qm, f, scale);
      b[j + 12] = quant(x3, j + 12, qm, f, scale);
    } else {
      b[j] = (int16_t)x0;
      b[j + 8] = (int16_t)x2;
      b[j + 4] = (int16_t)x1;
      b[j + 12] = (int16_t)x3;
    }
  }
}

struct Dq {
  int qp, f, rs, lsh, qm;
};
ICU_HD Dq dq_setup(int qp) {
  Dq d;
  d.qp = qp;
  d.qm = qp % 6;
  const int qe = qp / 6;
  d.f = qe <= 3 ? 1 << (3 - qe) : 0;
  d.rs = 4 - qe;
  d.lsh = qe - 4;
  return d;
}
ICU_HD int dq(const Dq& d, int lvl, int p, const int* w) {
  const int x = lvl * (kV[d.qm][pos_class(p)] * (w ? w[p] : 16));
  return d.qp < 24 ? (x + d.f) >> d.rs : x * (1 << d.lsh);
}

#define ICU_INV_BFLY(a0, a1, a2, a3, y0, y1, y2, y3) \
  do {                                               \
    const int s0_ = (a0) + (a2), s1_ = (a0) - (a2);  \
    const int s2_ = ((a1) >> 1) - (a3), s3_ = (a1) + ((a3) >> 1); \
    y0 = s0_ + s3_; y1 = s1_ + s2_; y2 = s1_ - s2_; y3 = s0_ - s3_; \
  } while (0)

ICU_HD void inv4x4(int16_t (&b)[16], int qp, int mode, int cols_first,
                                       const int* w) {
  const Dq d = dq_setup(qp);
  if (mode == 2) {
#pragma unroll
    for (int p = 0; p < 16; p++) b[p] = (int16_t)dq(d, b[p], p, w);
    return;
  }
  const bool deq = mode == 0;
  if (!cols_first) {
#pragma unroll
    for (int j = 0; j < 16; j += 4) {
      int x0 = b[j], x1 = b[j + 1], x2 = b[j + 2], x3 = b[j + 3];
      if (deq) { x0 = dq(d, x0, j, w); x1 = dq(d, x1, j + 1, w); x2 = dq(d, x2, j + 2, w); x3 = dq(d, x3, j + 3, w); }
      int y0, y1, y2, y3;
      ICU_INV_BFLY(x0, x1, x2, x3, y0, y1, y2, y3);
      b[j] = (int16_t)y0; b[j + 1] = (int16_t)y1; b[j + 2] = (int16_t)y2; b[j + 3] = (int16_t)y3;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int y0, y1, y2, y3;
      ICU_INV_BFLY((int)b[j], (int)b[j + 4], (int)b[j + 8], (int)b[j + 12], y0, y1, y2, y3);
      b[j] = (int16_t)((y0 + 32) >> 6); b[j + 4] = (int16_t)((y1 + 32) >> 6);
      b[j + 8] = (int16_t)((y2 + 32) >> 6); b[j + 12] = (int16_t)((y3 + 32) >> 6);
    }
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int x0 = b[j], x1 = b[j + 4], x2 = b[j + 8], x3 = b[j + 12];
      if (deq) { x0 = dq(d, x0, j, w); x1 = dq(d, x1, j + 4, w); x2 = dq(d, x2, j + 8, w); x3 = dq(d, x3, j + 12, w); }
      int y0, y1, y2, y3;
      ICU_INV_BFLY(x0, x1, x2, x3, y0, y1, y2, y3);
      b[j] = (int16_t)y0; b[j + 4] = (int16_t)y1; b[j + 8] = (int16_t)y2; b[j + 12] = (int16_t)y3;
    }
#pragma unroll
    for (int j = 0; j < 16; j += 4) {
      int y0, y1, y2, y3;
      ICU_INV_BFLY((int)b[j], (int)b[j + 1], (int)b[j + 2], (int)b[j + 3], y0, y1, y2, y3);
      b[j] = (int16_t)((y0 + 32) >> 6); b[j + 1] = (int16_t)((y1 + 32) >> 6);
      b[j + 2] = (int16_t)((y2 + 32) >> 6); b[j + 3] = (int16_t)((y3 + 32) >> 6);
    }
  }
}

// Element offset of coefficient p of 4x4 block `blk` in the given layout.
ICU_HD long elem_off(long blk, int p, int layout) {
  if (layout == 0) return blk * 16 + p;
  const long mb = blk >> 4;
  const int bi = (int)(blk >> 2) & 3, bj = (int)blk & 3;
  return mb * 256 + (4 * bi + (p >> 2)) * 16 + 4 * bj + (p & 3);
}

}  // namespace tx
}  // namespace icu
