/*
 * CPU oracle of the multi-reference, variable-block-size integer motion
 * search (16x16 / 16x8 / 8x16 / 8x8 from 4x4 SADs, rate term) and of its
 * motion-compensated prediction.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke entry point and the
 * benchmark's CPU baseline as the checker, never by the product library.
 *
 * The reference has no multi-reference or partitioned estimator; this
 * search composes its pieces:
 *   - 4x4 / 8x8 / 16x16 sums of absolute differences as
 *     OverlayMem2Dv2::Tad4x4 / Tad8x8 / Tad16x16
 *     (src/ImageUtils/OverlayMem2Dv2.cpp:2555-2735);
 *   - the window of each MB as GetMotionRange with half-width R
 *     (src/CodecUtils/MotionEstimatorH264ImplFull.cpp:1029-1081, the
 *     exhaustive +-R form), every partition of an MB sharing its window;
 *   - reference samples outside the picture as the edge-replicating padded
 *     overlay (src/ImageUtils/OverlayExtMem2Dv2.cpp:86-217).
 * Decision rule (no reference counterpart), written the slow, obvious way:
 *   J(p, r, v) = SAD_p(r, v) + lambda * (|vx| + |vy|), full-pel v;
 *   per partition the least J wins, ties to the lower reference index,
 *   then to the earlier candidate in raster order (rows, then columns);
 *   the MB mode is the least sum of its partitions' J, ties to the mode
 *   with fewer partitions (16x16, 16x8, 8x16, 8x8).
 */
#include <limits.h>
#include <stdlib.h>

#include "oracle.h"

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static int px(const uint8_t *p, int W, int H, int x, int y) {
  return p[(long)clampi(y, 0, H - 1) * W + clampi(x, 0, W - 1)];
}

/* SAD of the 4x4 block at (bx, by) of the MB at (n, m) against reference
 * position (n + bx + vx, m + by + vy). */
static int sad4x4(const uint8_t *cur, const uint8_t *ref, int W, int H, int n, int m, int bx,
                  int by, int vx, int vy) {
  int acc = 0;
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) {
      const int x = n + bx + c, y = m + by + r;
      acc += abs((int)cur[(long)y * W + x] - px(ref, W, H, x + vx, y + vy));
    }
  return acc;
}

/* Partition p of an MB: 0 16x16; 1, 2 16x8 top / bottom; 3, 4 8x16 left /
 * right; 5..8 8x8 in raster order.  Returns whether 8x8 quadrant k
 * (0 top-left, 1 top-right, 2 bottom-left, 3 bottom-right) belongs to p. */
static int part_has(int p, int k) {
  switch (p) {
    case 0: return 1;
    case 1: return k < 2;
    case 2: return k >= 2;
    case 3: return (k & 1) == 0;
    case 4: return (k & 1) == 1;
    default: return k == p - 5;
  }
}

int orc_vbs_estimate(const uint8_t *cur, const uint8_t *refs, int nrefs, int W, int H, int R,
                     int lambda, uint8_t *mode_out, int16_t *mv8_out, uint8_t *ref8_out,
                     int32_t *cost_out) {
  if (W % 16 || H % 16 || nrefs < 1 || R < 1) return -1;
  const int mbw = W / 16, mbh = H / 16, bw = W / 8;
  for (int mr = 0; mr < mbh; mr++)
    for (int mc = 0; mc < mbw; mc++) {
      const int n = 16 * mc, m = 16 * mr;
      /* GetMotionRange, half-width R, 16-pel border */
      const int xl = (n - R >= -16) ? -R : -(n + 16);
      const int xr = (n + R < W) ? R : W - n;
      const int yu = (m - R >= -16) ? -R : -(m + 16);
      const int yd = (m + R < H) ? R : H - m;
      long bestJ[9];
      int bestR[9], bestX[9], bestY[9];
      for (int p = 0; p < 9; p++) bestJ[p] = LONG_MAX;
      for (int r = 0; r < nrefs; r++) {
        const uint8_t *ref = refs + (long)r * W * H;
        for (int vy = yu; vy <= yd; vy++)
          for (int vx = xl; vx <= xr; vx++) {
            int s4[4][4];
            for (int bi = 0; bi < 4; bi++)
              for (int bj = 0; bj < 4; bj++) s4[bi][bj] = sad4x4(cur, ref, W, H, n, m, 4 * bj, 4 * bi, vx, vy);
            int q[4];
            for (int k = 0; k < 4; k++) {
              const int i0 = 2 * (k >> 1), j0 = 2 * (k & 1);
              q[k] = s4[i0][j0] + s4[i0][j0 + 1] + s4[i0 + 1][j0] + s4[i0 + 1][j0 + 1];
            }
            const long rate = (long)lambda * (abs(vx) + abs(vy));
            for (int p = 0; p < 9; p++) {
              long s = 0;
              for (int k = 0; k < 4; k++)
                if (part_has(p, k)) s += q[k];
              const long J = s + rate;
              if (J < bestJ[p]) { /* strict: earlier (ref, row, column) keeps ties */
                bestJ[p] = J;
                bestR[p] = r;
                bestX[p] = vx;
                bestY[p] = vy;
              }
            }
          }
      }
      const long modeJ[4] = {bestJ[0], bestJ[1] + bestJ[2], bestJ[3] + bestJ[4],
                             bestJ[5] + bestJ[6] + bestJ[7] + bestJ[8]};
      int mode = 0;
      for (int k = 1; k < 4; k++)
        if (modeJ[k] < modeJ[mode]) mode = k;
      const int mb = mr * mbw + mc;
      mode_out[mb] = (uint8_t)mode;
      cost_out[mb] = (int32_t)modeJ[mode];
      for (int k = 0; k < 4; k++) {
        int p = 0;
        if (mode == 1) p = k < 2 ? 1 : 2;
        if (mode == 2) p = (k & 1) ? 4 : 3;
        if (mode == 3) p = 5 + k;
        const long b = (long)(2 * mr + (k >> 1)) * bw + 2 * mc + (k & 1);
        mv8_out[2 * b] = (int16_t)(4 * bestX[p]);
        mv8_out[2 * b + 1] = (int16_t)(4 * bestY[p]);
        ref8_out[b] = (uint8_t)bestR[p];
      }
    }
  return 0;
}

/* Prediction of every 8x8 block: the block of reference ref8[b] at the
 * quarter-pel vector mv8[b] floored to full pels (an arithmetic shift right
 * by 2), edge-replicated outside the picture
 * (MotionCompensatorH264ImplStd::Compensate at an integer vector,
 * src/CodecUtils/MotionCompensatorH264ImplStd.cpp:410-437). */
void orc_vbs_pred(const uint8_t *refs, int W, int H, const int16_t *mv8, const uint8_t *ref8,
                  int16_t *pred_out) {
  const int bw = W / 8;
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const long b = (long)(y / 8) * bw + x / 8;
      const uint8_t *ref = refs + (long)ref8[b] * W * H;
      const int fx = mv8[2 * b] >= 0 ? mv8[2 * b] / 4 : -((-mv8[2 * b] + 3) / 4);
      const int fy = mv8[2 * b + 1] >= 0 ? mv8[2 * b + 1] / 4 : -((-mv8[2 * b + 1] + 3) / 4);
      pred_out[(long)y * W + x] = (int16_t)px(ref, W, H, x + fx, y + fy);
    }
}
