/*
 * CPU restatement of the reference's H.264 4x4 integer transforms with
 * quantisation / dequantisation, luma motion compensation and residual
 * coding.  TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path.
 *
 * Follows:
 *   src/CodecUtils/FastForward4x4ITImpl1.cpp:57-212,223-278  (rows then cols,
 *       MF table :57-65, position classes :67-73, f/scale :269-278)
 *   src/CodecUtils/FastForward4x4On16x16ITImpl1.cpp:457-755  (same per 4x4 of
 *       a 16x16 raster MB)
 *   src/CodecUtils/FastInverse4x4ITImpl1.cpp:60-236,251-302    (rows then cols)
 *   src/CodecUtils/FastInverse4x4On16x16ITImpl1.cpp:120-360   (cols then rows)
 *   src/ImageUtils/OverlayMem2Dv2.cpp:1195-1220,1343-1368      (Sub16x16 and
 *       Add16x16WithClip255)
 *   src/CodecUtils/MotionCompensatorH264ImplStd.cpp:410-437    (luma compensate)
 *
 * Every (short) store of the reference is an int16_t store here, at the same
 * point.  Intermediates are int; inputs are expected in the range where the
 * reference's int arithmetic does not overflow (|x| <= 4551 for the forward
 * IT+Q, |level| * 29 * 16 << 4 within int for the inverse), which covers
 * 8-bit residuals and every level they quantise to.
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

#include "oracle.h"

static const int FWD_MF[6][3] = {{13107, 5243, 8066}, {11916, 4660, 7490},
                                 {10082, 4194, 6554}, {9362, 3647, 5825},
                                 {8192, 3355, 5243},  {7282, 2893, 4559}};
static const int INV_V[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16},
                                {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
/* 0 = (even,even), 1 = (odd,odd), 2 = mixed, by raster position in the 4x4 */
static const int POS_CLASS[16] = {0, 2, 0, 2, 2, 1, 2, 1, 0, 2, 0, 2, 2, 1, 2, 1};

static inline int16_t quant(int x, int r, int qm, int f, int scale) {
  if (x < 0) return (int16_t)(-((((-x) * FWD_MF[qm][r]) + f) >> scale));
  return (int16_t)(((x * FWD_MF[qm][r]) + f) >> scale);
}

void orc_fwd4x4(int16_t *b, int qp, int intra, int mode) {
  const int qm = qp % 6, qe = qp / 6, scale = 15 + qe;
  const int f = intra ? (1 << (15 + qe)) / 3 : (1 << (15 + qe)) / 6;
  if (mode == 2) {
    for (int p = 0; p < 16; p++) b[p] = quant(b[p], POS_CLASS[p], qm, f, scale);
    return;
  }
  for (int j = 0; j < 16; j += 4) {
    int s0 = b[j] + b[j + 3], s3 = b[j] - b[j + 3];
    int s1 = b[j + 1] + b[j + 2], s2 = b[j + 1] - b[j + 2];
    b[j] = (int16_t)(s0 + s1);
    b[j + 2] = (int16_t)(s0 - s1);
    b[j + 1] = (int16_t)(s2 + s3 * 2);
    b[j + 3] = (int16_t)(s3 - s2 * 2);
  }
  for (int j = 0; j < 4; j++) {
    int s0 = b[j] + b[j + 12], s3 = b[j] - b[j + 12];
    int s1 = b[j + 4] + b[j + 8], s2 = b[j + 4] - b[j + 8];
    int x0 = s0 + s1, x2 = s0 - s1, x1 = s2 + s3 * 2, x3 = s3 - s2 * 2;
    if (mode == 0) {
      b[j] = quant(x0, POS_CLASS[j], qm, f, scale);
      b[j + 8] = quant(x2, POS_CLASS[j + 8], qm, f, scale);
      b[j + 4] = quant(x1, POS_CLASS[j + 4], qm, f, scale);
      b[j + 12] = quant(x3, POS_CLASS[j + 12], qm, f, scale);
    } else {
      b[j] = (int16_t)x0;
      b[j + 8] = (int16_t)x2;
      b[j + 4] = (int16_t)x1;
      b[j + 12] = (int16_t)x3;
    }
  }
}

typedef struct { int ls[16], qp, f, rs, lsh; } dq_t;

static void dq_setup(dq_t *d, int qp, const int *w) {
  const int qm = qp % 6, qe = qp / 6;
  for (int p = 0; p < 16; p++) d->ls[p] = INV_V[qm][POS_CLASS[p]] * (w ? w[p] : 16);
  d->qp = qp;
  d->f = qe <= 3 ? 1 << (3 - qe) : 0;
  d->rs = 4 - qe;
  d->lsh = qe - 4;
}

static inline int dq(const dq_t *d, int lvl, int p) {
  int x = lvl * d->ls[p];
  return d->qp < 24 ? (x + d->f) >> d->rs : x * (1 << d->lsh);
}

/* 1-D inverse butterfly on (a0,a1,a2,a3) -> y0..y3, no rounding. */
#define INV_BFLY(a0, a1, a2, a3, y0, y1, y2, y3) \
  do {                                            \
    int s0_ = (a0) + (a2), s1_ = (a0) - (a2);     \
    int s2_ = ((a1) >> 1) - (a3), s3_ = (a1) + ((a3) >> 1); \
    y0 = s0_ + s3_; y1 = s1_ + s2_; y2 = s1_ - s2_; y3 = s0_ - s3_; \
  } while (0)

static void inv_impl(int16_t *b, int qp, int mode, const int *w, int cols_first) {
  dq_t d;
  dq_setup(&d, qp, w);
  if (mode == 2) {
    for (int p = 0; p < 16; p++) b[p] = (int16_t)dq(&d, b[p], p);
    return;
  }
  const int deq = (mode == 0);
  if (!cols_first) {
    for (int j = 0; j < 16; j += 4) {
      int x0 = b[j], x1 = b[j + 1], x2 = b[j + 2], x3 = b[j + 3];
      if (deq) { x0 = dq(&d, x0, j); x1 = dq(&d, x1, j + 1); x2 = dq(&d, x2, j + 2); x3 = dq(&d, x3, j + 3); }
      int y0, y1, y2, y3;
      INV_BFLY(x0, x1, x2, x3, y0, y1, y2, y3);
      b[j] = (int16_t)y0; b[j + 1] = (int16_t)y1; b[j + 2] = (int16_t)y2; b[j + 3] = (int16_t)y3;
    }
    for (int j = 0; j < 4; j++) {
      int y0, y1, y2, y3;
      INV_BFLY(b[j], b[j + 4], b[j + 8], b[j + 12], y0, y1, y2, y3);
      b[j] = (int16_t)((y0 + 32) >> 6); b[j + 4] = (int16_t)((y1 + 32) >> 6);
      b[j + 8] = (int16_t)((y2 + 32) >> 6); b[j + 12] = (int16_t)((y3 + 32) >> 6);
    }
  } else {
    for (int j = 0; j < 4; j++) {
      int x0 = b[j], x1 = b[j + 4], x2 = b[j + 8], x3 = b[j + 12];
      if (deq) { x0 = dq(&d, x0, j); x1 = dq(&d, x1, j + 4); x2 = dq(&d, x2, j + 8); x3 = dq(&d, x3, j + 12); }
      int y0, y1, y2, y3;
      INV_BFLY(x0, x1, x2, x3, y0, y1, y2, y3);
      b[j] = (int16_t)y0; b[j + 4] = (int16_t)y1; b[j + 8] = (int16_t)y2; b[j + 12] = (int16_t)y3;
    }
    for (int j = 0; j < 16; j += 4) {
      int y0, y1, y2, y3;
      INV_BFLY(b[j], b[j + 1], b[j + 2], b[j + 3], y0, y1, y2, y3);
      b[j] = (int16_t)((y0 + 32) >> 6); b[j + 1] = (int16_t)((y1 + 32) >> 6);
      b[j + 2] = (int16_t)((y2 + 32) >> 6); b[j + 3] = (int16_t)((y3 + 32) >> 6);
    }
  }
}

void orc_inv4x4(int16_t *b, int qp, int mode, const int *w) { inv_impl(b, qp, mode, w, 0); }
void orc_inv4x4_colsfirst(int16_t *b, int qp, int mode, const int *w) { inv_impl(b, qp, mode, w, 1); }

void orc_fwd4x4_n(int16_t *blks, long n, int qp, int intra, int mode) {
  for (long i = 0; i < n; i++) orc_fwd4x4(blks + 16 * i, qp, intra, mode);
}
void orc_inv4x4_n(int16_t *blks, long n, int qp, int mode, int cols_first) {
  for (long i = 0; i < n; i++) inv_impl(blks + 16 * i, qp, mode, NULL, cols_first);
}

static void mb_gather(const int16_t *mb, int bi, int bj, int16_t *b) {
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) b[r * 4 + c] = mb[(4 * bi + r) * 16 + 4 * bj + c];
}
static void mb_scatter(int16_t *mb, int bi, int bj, const int16_t *b) {
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) mb[(4 * bi + r) * 16 + 4 * bj + c] = b[r * 4 + c];
}

void orc_fwd_mb_n(int16_t *mbs, long n, int qp, int intra, int mode) {
  int16_t b[16];
  for (long i = 0; i < n; i++)
    for (int bi = 0; bi < 4; bi++)
      for (int bj = 0; bj < 4; bj++) {
        mb_gather(mbs + 256 * i, bi, bj, b);
        orc_fwd4x4(b, qp, intra, mode);
        mb_scatter(mbs + 256 * i, bi, bj, b);
      }
}

void orc_inv_mb_n(int16_t *mbs, long n, int qp, int mode, int cols_first) {
  int16_t b[16];
  for (long i = 0; i < n; i++)
    for (int bi = 0; bi < 4; bi++)
      for (int bj = 0; bj < 4; bj++) {
        mb_gather(mbs + 256 * i, bi, bj, b);
        inv_impl(b, qp, mode, NULL, cols_first);
        mb_scatter(mbs + 256 * i, bi, bj, b);
      }
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

void orc_mc_luma(const int16_t *ref, int W, int H, const int16_t *mv, int16_t *pred) {
  int blk[256];
  const int mbw = W / 16, mbh = H / 16;
  for (int row = 0; row < mbh; row++)
    for (int col = 0; col < mbw; col++) {
      const int n = col * 16, m = row * 16, v = row * mbw + col;
      const int mvx = mv[2 * v], mvy = mv[2 * v + 1];
      const int fx = mvx / 4, fy = mvy / 4, qx = mvx % 4, qy = mvy % 4;
      if (!qx && !qy) {
        for (int r = 0; r < 16; r++)
          for (int c = 0; c < 16; c++)
            blk[r * 16 + c] = ref[(long)clampi(m + fy + r, 0, H - 1) * W + clampi(n + fx + c, 0, W - 1)];
      } else {
        orc_quarter_read(ref, W, H, n + fx, m + fy, qx, qy, blk);
      }
      for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++) pred[(long)(m + r) * W + n + c] = (int16_t)blk[r * 16 + c];
    }
}

void orc_code_residual(const int16_t *cur, const int16_t *pred, int W, int H, int qp,
                       int intra, int cols_first, int16_t *levels, int16_t *recon) {
  int16_t b[16];
  for (int by = 0; by < H; by += 4)
    for (int bx = 0; bx < W; bx += 4) {
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
          long o = (long)(by + r) * W + bx + c;
          b[r * 4 + c] = (int16_t)(cur[o] - pred[o]);
        }
      orc_fwd4x4(b, qp, intra, 0);
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) levels[(long)(by + r) * W + bx + c] = b[r * 4 + c];
      inv_impl(b, qp, 0, NULL, cols_first);
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
          long o = (long)(by + r) * W + bx + c;
          recon[o] = (int16_t)clampi(pred[o] + b[r * 4 + c], 0, 255);
        }
    }
}
