/*
 * CPU restatement of the reference's H.264 16x16 block-matching motion
 * estimator.  TEST INFRASTRUCTURE ONLY: this file is the parity checker for
 * the HIP path.  Only tests/, the smoke entry point and the CPU-baseline leg
 * of the benchmark may load it; no product code links or calls it.
 *
 * It follows, step for step:
 *   src/CodecUtils/MotionEstimatorH264ImplFull.cpp:330-611   Estimate()
 *   src/CodecUtils/MotionEstimatorH264ImplFull.cpp:1029-1081 GetMotionRange (9-arg)
 *   src/CodecUtils/MotionEstimatorH264ImplFull.cpp:1095-1255 half/quarter window
 *   include/CodecUtils/MotionEstimatorH264ImplFull.h:140-168 GetBestVector
 *   include/GeneralUtils/Fifo.h:54-106                       candidate fifo
 *   src/ImageUtils/OverlayMem2Dv2.cpp:704-1052               QuarterRead
 *   src/ImageUtils/OverlayMem2Dv2.cpp:1623-1983,2655-2983    Tsd/Tad 16x16
 *   src/ImageUtils/OverlayExtMem2Dv2.cpp:86-217              padded reference
 *   src/CodecUtils/MacroBlockH264.cpp:757-816,923-936        median predictor
 *
 * The padded reference (border 19, edge replication) is restated as a
 * coordinate clamp; every access the estimator makes lies inside the padded
 * area ([-19, W+19) x [-19, H+19)), which the accessor asserts.
 *
 * Switches beyond the as-shipped class (all default to the shipped value):
 *   metric   0 = SSD (shipped, Full.cpp:74 #undef MEH264IF_ABS_DIFF), 1 = SAD
 *   window   0 = reference rule +-(R-1) (Full.cpp:399 -> :1029), 1 = +-R
 *   filter   1 = running-best candidate filter (Full.cpp:410-411), 0 = off
 *   subpel   2 = half+quarter (shipped), 1 = half only, 0 = integer only
 *   pred     1 = H.264 median of A/B/C(D) (H264MotionVectorPredictorImpl1),
 *            0 = zero predictor
 *   bestvec  1 = GetBestVector 3-way choice (shipped), 0 = keep refined best
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdio.h>

#include "oracle.h"

#define PAD 19 /* 16 + MEH264IF_PADDING (Full.cpp:244) */

typedef struct {
  const int16_t *p;
  int w, h;
} plane_t;

static inline int pix(const plane_t *pl, int x, int y) {
  if (x < -PAD || x >= pl->w + PAD || y < -PAD || y >= pl->h + PAD) {
    fprintf(stderr, "oracle: access (%d,%d) outside padded reference\n", x, y);
    abort();
  }
  if (x < 0) x = 0;
  if (x > pl->w - 1) x = pl->w - 1;
  if (y < 0) y = 0;
  if (y > pl->h - 1) y = pl->h - 1;
  return pl->p[(long)y * pl->w + x];
}

static inline int clip255(int x) { return x < 0 ? 0 : (x > 255 ? 255 : x); }

/* Block distortion between the 16x16 input block at (n,m) and a 16x16 block
 * of samples blk[256] (row-major).  Tsd16x16 / Tad16x16 semantics; the
 * LessThan early exit only ever returns a partial sum > min, which is never
 * accepted, so the exact sum is equivalent for every decision. */
static int blk_dist(const int16_t *cur, int W, int n, int m, const int *blk,
                    int metric) {
  int acc = 0;
  for (int r = 0; r < 16; r++)
    for (int c = 0; c < 16; c++) {
      int d = cur[(long)(m + r) * W + n + c] - blk[r * 16 + c];
      acc += metric ? (d < 0 ? -d : d) : d * d;
    }
  return acc;
}

static void read_full(const plane_t *ref, int x, int y, int *blk) {
  for (int r = 0; r < 16; r++)
    for (int c = 0; c < 16; c++) blk[r * 16 + c] = pix(ref, x + c, y + r);
}

/* H.264 6-tap (1,-5,20,20,-5,1) helpers (OverlayMem2Dv2.cpp:92-98). */
#define TAP6(a, b, c, d, e, f) ((a) - 5 * (b) + 20 * (c) + 20 * (d) - 5 * (e) + (f))
static inline int vtap(const plane_t *p, int x, int y) {
  return TAP6(pix(p, x, y - 2), pix(p, x, y - 1), pix(p, x, y), pix(p, x, y + 1),
              pix(p, x, y + 2), pix(p, x, y + 3));
}
static inline int htap(const plane_t *p, int x, int y) {
  return TAP6(pix(p, x - 2, y), pix(p, x - 1, y), pix(p, x, y), pix(p, x + 1, y),
              pix(p, x + 2, y), pix(p, x + 3, y));
}
static inline int get_b(const plane_t *p, int x, int y) { return clip255((htap(p, x, y) + 16) >> 5); }
static inline int get_h(const plane_t *p, int x, int y) { return clip255((vtap(p, x, y) + 16) >> 5); }
static inline int get_j(const plane_t *p, int x, int y) {
  int s = TAP6(vtap(p, x - 2, y), vtap(p, x - 1, y), vtap(p, x, y), vtap(p, x + 1, y),
               vtap(p, x + 2, y), vtap(p, x + 3, y));
  return clip255((s + 512) >> 10);
}

/* One quarter-pel sample at full-pel (x,y) with fractions fx,fy in 0..3
 * (OverlayMem2Dv2::QuarterRead switch, :729-1010; H.264 8.4.2.2.1). */
int orc_qpel_sample(const int16_t *ref, int W, int H, int x, int y, int fx, int fy) {
  plane_t p = {ref, W, H};
  int G = pix(&p, x, y);
  switch (fx | (fy << 2)) {
    case 0: return G;
    case 2: return get_b(&p, x, y);
    case 8: return get_h(&p, x, y);
    case 10: return get_j(&p, x, y);
    case 1: return (G + get_b(&p, x, y) + 1) >> 1;                       /* a */
    case 3: return (pix(&p, x + 1, y) + get_b(&p, x, y) + 1) >> 1;       /* c */
    case 4: return (G + get_h(&p, x, y) + 1) >> 1;                       /* d */
    case 12: return (pix(&p, x, y + 1) + get_h(&p, x, y) + 1) >> 1;      /* n */
    case 6: return (get_j(&p, x, y) + get_b(&p, x, y) + 1) >> 1;         /* f */
    case 14: return (get_j(&p, x, y) + get_b(&p, x, y + 1) + 1) >> 1;    /* q */
    case 9: return (get_j(&p, x, y) + get_h(&p, x, y) + 1) >> 1;         /* i */
    case 11: return (get_j(&p, x, y) + get_h(&p, x + 1, y) + 1) >> 1;    /* k */
    case 5: return (get_b(&p, x, y) + get_h(&p, x, y) + 1) >> 1;         /* e */
    case 7: return (get_b(&p, x, y) + get_h(&p, x + 1, y) + 1) >> 1;     /* g */
    case 13: return (get_h(&p, x, y) + get_b(&p, x, y + 1) + 1) >> 1;    /* p */
    default: return (get_h(&p, x + 1, y) + get_b(&p, x, y + 1) + 1) >> 1; /* r (15) */
  }
}

/* QuarterRead of a 16x16 block whose origin is full-pel (x,y) with signed
 * quarter offsets qx,qy in -3..3: negative fractions are reflected to the
 * full-pel position to the left/above (OverlayMem2Dv2.cpp:709-724). */
void orc_quarter_read(const int16_t *ref, int W, int H, int x, int y, int qx, int qy,
                      int *blk) {
  int fox = 0, fx = qx, foy = 0, fy = qy;
  if (qx < 0) { fox = -1; fx = 4 + qx; }
  if (qy < 0) { foy = -1; fy = 4 + qy; }
  for (int r = 0; r < 16; r++)
    for (int c = 0; c < 16; c++)
      blk[r * 16 + c] = orc_qpel_sample(ref, W, H, x + c + fox, y + r + foy, fx, fy);
}

/* GetMotionRange 9-arg with zero offset (Full.cpp:1029-1081), or the
 * exhaustive +-R variant that keeps the same picture-border clamp. */
static void motion_range(int n, int m, int W, int H, int R, int window, int *xl,
                         int *xr, int *yu, int *yd) {
  int L = window ? R : R - 1;
  const int boundary = 16; /* _extBoundary - MEH264IF_PADDING */
  *xl = (n - L >= -boundary) ? -L : -(n + boundary);
  *xr = (n + L < W) ? L : W - n;
  *yu = (m - L >= -boundary) ? -L : -(m + boundary);
  *yd = (m + L < H) ? L : H - m;
}

/* Median of three (MacroBlockH264.cpp:923-936). */
static int median3(int x, int y, int z) {
  int mn, mx;
  if (y - z < 0) { mn = y; mx = z; } else { mn = z; mx = y; }
  if (x < mn) mn = x;
  if (x > mx) mx = x;
  return x + y + z - mn - mx;
}

/* 16x16 median prediction for MB (col,row) from MBs already decided in this
 * Estimate call (MacroBlockH264.cpp:757-816 with neighbours set up by
 * MacroBlockH264::Initialise, :346-373, one slice). */
void orc_median_pred(const int16_t *mv, int mbw, int col, int row, int *px, int *py) {
  int Ax = 0, Ay = 0, Bx = 0, By = 0, Cx = 0, Cy = 0;
  int hasB = row > 0, hasC = row > 0 && col + 1 < mbw;
  if (col > 0) { const int16_t *a = mv + 2 * (row * mbw + col - 1); Ax = a[0]; Ay = a[1]; }
  if (hasB) { const int16_t *b = mv + 2 * ((row - 1) * mbw + col); Bx = b[0]; By = b[1]; }
  if (hasC) {
    const int16_t *c = mv + 2 * ((row - 1) * mbw + col + 1); Cx = c[0]; Cy = c[1];
  } else if (col > 0 && row > 0) {
    const int16_t *d = mv + 2 * ((row - 1) * mbw + col - 1); Cx = d[0]; Cy = d[1];
  }
  if (!hasB && !hasC) { Bx = Ax; By = Ay; Cx = Ax; Cy = Ay; }
  *px = median3(Ax, Bx, Cx);
  *py = median3(Ay, By, Cy);
}

/* GetBestVector (Full.h:140-168) over the fifo [best, pred, zero, zero, zero]
 * (Fifo::AddFirstIn on an empty fifo fills every slot, Fifo.h:72-92).
 * Returns the fifo position chosen. */
static int best_vector(const int *xs, const int *ys, const int *ds, int xref, int yref,
                       int *mx, int *my) {
  int pos = 0;
  double cm = (double)((xs[0] - xref) * (xs[0] - xref) + (ys[0] - yref) * (ys[0] - yref));
  double cc = ((double)ds[0] / 256.0) + (1.05 * sqrt(cm));
  *mx = xs[0];
  *my = ys[0];
  for (int i = 1; i < 5; i++) {
    double nm = (double)((xs[i] - xref) * (xs[i] - xref) + (ys[i] - yref) * (ys[i] - yref));
    double nc = ((double)ds[i] / 256.0) + (1.05 * sqrt(nm));
    if (nc < cc) { *mx = xs[i]; *my = ys[i]; cc = nc; pos = i; }
  }
  return pos;
}

/* Half-pel window values (LoadHalfQuartPelWindow, Full.cpp:1095-1182) around
 * full-pel origin (ox,oy), on the estimator's 88x88 quarter-pel grid with
 * full pel (3,3) at index (12,12).  Only the even/even positions used by the
 * stages below are produced; they equal the H.264 G/b/h/j samples. */
typedef struct { int16_t w[88][88]; } qwin_t;

static void load_half_window(qwin_t *W8, const plane_t *ref, int ox, int oy) {
  /* positions 8..76 even x even, in 1/4-pel units, relative to full pel (0,0)
     of the window which is ref (ox-3, oy-3) */
  for (int qr = 8; qr <= 76; qr += 2)
    for (int qc = 8; qc <= 76; qc += 2) {
      int fr = qr >> 2, fc = qc >> 2;
      int x = ox - 3 + fc, y = oy - 3 + fr;
      int hr = (qr & 2) != 0, hc = (qc & 2) != 0;
      int v;
      if (!hr && !hc) v = pix(ref, x, y);
      else if (!hr && hc) v = get_b(ref, x, y);
      else if (hr && !hc) v = get_h(ref, x, y);
      else v = get_j(ref, x, y);
      W8->w[qr][qc] = (int16_t)v;
    }
}

/* LoadQuartPelWindow (Full.cpp:1198-1222): the 8 quarter positions around
 * the winning half-pel position, each the rounded mean of two window
 * samples exactly as the reference pairs them. */
static void load_quarter_window(qwin_t *W8, int hx, int hy) {
  for (int r = 0; r < 16; r++)
    for (int c = 0; c < 16; c++) {
      int qr = 12 + hy + 4 * r, qc = 12 + hx + 4 * c;
      int16_t(*w)[88] = W8->w;
      w[qr - 1][qc - 1] = (int16_t)((w[qr - 2][qc - 2] + w[qr][qc] + 1) >> 1);
      w[qr - 1][qc] = (int16_t)((w[qr - 2][qc] + w[qr][qc] + 1) >> 1);
      w[qr - 1][qc + 1] = (int16_t)((w[qr - 2][qc] + w[qr][qc + 2] + 1) >> 1);
      w[qr][qc - 1] = (int16_t)((w[qr][qc - 2] + w[qr][qc] + 1) >> 1);
      w[qr][qc + 1] = (int16_t)((w[qr][qc + 2] + w[qr][qc] + 1) >> 1);
      w[qr + 1][qc - 1] = (int16_t)((w[qr][qc - 2] + w[qr + 2][qc] + 1) >> 1);
      w[qr + 1][qc] = (int16_t)((w[qr + 2][qc] + w[qr][qc] + 1) >> 1);
      w[qr + 1][qc + 1] = (int16_t)((w[qr + 2][qc] + w[qr][qc + 2] + 1) >> 1);
    }
}

static void window_read(const qwin_t *W8, int qx, int qy, int *blk) {
  for (int r = 0; r < 16; r++)
    for (int c = 0; c < 16; c++) blk[r * 16 + c] = W8->w[12 + qy + 4 * r][12 + qx + 4 * c];
}

static const int SUBPOS[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {-1, 0},
                                 {1, 0},   {-1, 1}, {0, 1},  {1, 1}}; /* Full.cpp:77-81 */

int orc_me_estimate(const orc_me_cfg *cfg, const int16_t *cur, const int16_t *refp,
                    const uint8_t *include, int16_t *mv_out, int32_t *dist_out,
                    int32_t *int_out, long *avg_out) {
  return orc_me_estimate_subset(cfg, cur, refp, include, 1, mv_out, dist_out, int_out, avg_out);
}

/* As orc_me_estimate, but only every `step`-th MB in raster order is decided
 * (the others are left untouched).  Only meaningful with the zero predictor,
 * where MBs do not depend on each other. */
int orc_me_estimate_subset(const orc_me_cfg *cfg, const int16_t *cur, const int16_t *refp,
                           const uint8_t *include, int step, int16_t *mv_out, int32_t *dist_out,
                           int32_t *int_out, long *avg_out) {
  return orc_me_estimate_phase(cfg, cur, refp, include, step, 0, mv_out, dist_out, int_out, avg_out);
}

/* Only the MBs whose raster index is phase mod step (0 <= phase < step).
 * Disjoint phases write disjoint outputs, so calls over the phases of one
 * step may run on separate threads into the same arrays (avg_out NULL). */
int orc_me_estimate_phase(const orc_me_cfg *cfg, const int16_t *cur, const int16_t *refp,
                          const uint8_t *include, int step, int phase, int16_t *mv_out,
                          int32_t *dist_out, int32_t *int_out, long *avg_out) {
  if (step < 1 || (step > 1 && cfg->pred) || phase < 0 || phase >= step) return -3;
  const int W = cfg->width, H = cfg->height, R = cfg->range;
  if (W <= 0 || H <= 0 || (W & 15) || (H & 15) || R < 1) return -1;
  const int mbw = W / 16, mbh = H / 16;
  plane_t ref = {refp, W, H};
  qwin_t *win = (qwin_t *)malloc(sizeof(qwin_t));
  if (!win) return -2;
  memset(win, 0, sizeof(*win));
  int blk[256];
  long total = 0;
  int included = 0;

  for (int row = 0; row < mbh; row++)
    for (int col = 0; col < mbw; col++) {
      const int n = col * 16, m = row * 16, vec = row * mbw + col;
      if (vec % step != phase) continue;
      int predX = 0, predY = 0;
      if (cfg->pred) orc_median_pred(mv_out, mbw, col, row, &predX, &predY);
      int predXQ = predX % 4, predYQ = predY % 4;
      int predX0 = predX / 4, predY0 = predY / 4;
      int pXr = predX < 0 ? (predX - 2) / 4 : (predX + 2) / 4;
      int pYr = predY < 0 ? (predY - 2) / 4 : (predY + 2) / 4;

      read_full(&ref, n, m, blk);
      const int zeroD = blk_dist(cur, W, n, m, blk, cfg->metric);
      int minD = zeroD, mx = 0, my = 0;

      int xl, xr, yu, yd;
      motion_range(n, m, W, H, R, cfg->window, &xl, &xr, &yu, &yd);
      for (int i = yu; i <= yd; i++)
        for (int j = xl; j <= xr; j++) {
          if (!(i || j)) continue;
          if (cfg->filter &&
              ((i + my) < yu || (i + my) > yd || (j + mx) < xl || (j + mx) > xr))
            continue;
          read_full(&ref, n + j, m + i, blk);
          int d = blk_dist(cur, W, n, m, blk, cfg->metric);
          if (d <= minD) {
            if (d == minD) {
              int cx = mx - pXr, cy = my - pYr, nx = j - pXr, ny = i - pYr;
              if ((cy * cy + cx * cx) - (ny * ny + nx * nx) < 0) continue;
            }
            minD = d; mx = j; my = i;
          }
        }
      if (int_out) {
        int_out[3 * vec + 0] = mx;
        int_out[3 * vec + 1] = my;
        int_out[3 * vec + 2] = minD;
      }

      int mvx = mx * 4, mvy = my * 4, hmx = 0, hmy = 0, qmx = 0, qmy = 0;
      if (cfg->subpel >= 1) {
        load_half_window(win, &ref, n + mx, m + my);
        for (int k = 0; k < 8; k++) {
          int ox = 2 * SUBPOS[k][0], oy = 2 * SUBPOS[k][1];
          window_read(win, ox, oy, blk);
          int d = blk_dist(cur, W, n, m, blk, cfg->metric);
          if (d < minD) { minD = d; hmx = ox; hmy = oy; }
        }
        qmx = hmx; qmy = hmy;
        if (cfg->subpel >= 2) {
          load_quarter_window(win, hmx, hmy);
          for (int k = 0; k < 8; k++) {
            int ox = hmx + SUBPOS[k][0], oy = hmy + SUBPOS[k][1];
            window_read(win, ox, oy, blk);
            int d = blk_dist(cur, W, n, m, blk, cfg->metric);
            if (d < minD) { minD = d; qmx = ox; qmy = oy; }
          }
        }
      }
      mvx += qmx;
      mvy += qmy;

      int outx = mvx, outy = mvy, outd = minD;
      if (cfg->bestvec) {
        if (predX0 + n > W) predX0 = 16;
        if (predX0 + n < -16) predX0 = -16;
        if (predY0 + m > H) predY0 = 16;
        if (predY0 + m < -16) predY0 = -16;
        predX = predX0 * 4 + predXQ;
        predY = predY0 * 4 + predYQ;
        if (predXQ || predYQ)
          orc_quarter_read(refp, W, H, predX0 + n, predY0 + m, predXQ, predYQ, blk);
        else
          read_full(&ref, predX0 + n, predY0 + m, blk);
        int predD = blk_dist(cur, W, n, m, blk, cfg->metric);
        int xs[5] = {mvx, predX, 0, 0, 0}, ys[5] = {mvy, predY, 0, 0, 0};
        int ds[5] = {minD, predD, zeroD, zeroD, zeroD};
        int pos = best_vector(xs, ys, ds, predX, predY, &outx, &outy);
        outd = ds[pos];
      }
      if (include && include[vec]) { included++; total += outd; }
      mv_out[2 * vec] = (int16_t)outx;
      mv_out[2 * vec + 1] = (int16_t)outy;
      if (dist_out) dist_out[vec] = outd;
    }
  free(win);
  if (avg_out) *avg_out = included ? total / included : 0;
  return 0;
}

/* Tsd16x16 (metric 0) / Tad16x16 (metric 1) between two 16x16 blocks. */
int orc_block_dist(const int16_t *a256, const int16_t *b256, int metric) {
  int blk[256];
  for (int i = 0; i < 256; i++) blk[i] = b256[i];
  return blk_dist(a256, 16, 0, 0, blk, metric);
}

/* ------------------------------------------------------------------------
 * MotionEstimatorH264ImplUMHS::Estimate (src/CodecUtils/MotionEstimatorH264ImplUMHS.cpp:446-940),
 * SSD as shipped (:78 #undef MEH264IUMHS_ABS_DIFF), restated step for step:
 * starting points (predictor, zero, the previous frame's co-located vector),
 * early-termination tests, unsymmetrical cross, 5x5, uneven multi-hexagon,
 * extended hexagon and diamond loops, then the sub-pel stages shared with the
 * full-search estimator and the final test against the predictor.
 * The predictor is H264MotionVectorPredictorImpl1's distortion form
 * (MacroBlockH264::GetMbMotionMedianPred, MacroBlockH264.cpp:818-915),
 * including its comparison of the y median with C's x component (:893).
 * ---------------------------------------------------------------------- */

/* Pattern tables (UMHS.cpp:80-122). */
static const int U_SQ5[24][2] = {{-1, 0}, {1, 0},  {0, -1},  {0, 1},  {-1, -1}, {1, -1}, {-1, 1}, {1, 1},
                                 {0, -2}, {-2, 0}, {2, 0},   {0, 2},  {-1, -2}, {1, -2}, {-2, -1}, {2, -1},
                                 {-2, 1}, {2, 1},  {-1, 2},  {1, 2},  {-2, -2}, {2, -2}, {-2, 2}, {2, 2}};
static const int U_CROSS[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
static const int U_HEX16[16][2] = {{-4, 0}, {4, 0},  {-4, 1}, {4, 1},  {-4, -1}, {4, -1}, {-4, 2}, {4, 2},
                                   {-4, -2}, {4, -2}, {-2, 3}, {2, 3}, {-2, -3}, {2, -3}, {0, 4},  {0, -4}};
static const int U_EXTHEX[6][2] = {{-2, 0}, {-1, -2}, {1, -2}, {2, 0}, {1, 2}, {-1, 2}};

/* How often each branch ran since the last orc_umhs_stats call (so tests can
 * show that a set of inputs reaches every stage): 0 cross, 1 5x5, 2 hexagon,
 * 3 hexagon early break, 4 extended-hexagon moves, 5 diamond moves,
 * 6 previous-frame candidate tested, 7 predictor truncated, 8 predictor
 * chosen at the end, 9 exit to the diamond, 10 exit to the hexagon before
 * the cross, 11 exit after the cross, 12 exit after the 5x5. */
static long umhs_stat[13];
void orc_umhs_stats(long *out) {
  for (int i = 0; i < 13; i++) { out[i] = umhs_stat[i]; umhs_stat[i] = 0; }
}

/* MEH264IUMHS_COST (UMHS.cpp:167): distortion / 256 plus 1.05 x the
 * Euclidean distance to (xr, yr), truncated to int. */
static int umhs_cost(int d, int x, int y, int xr, int yr) {
  const double e = sqrt((double)((x - xr) * (x - xr) + (y - yr) * (y - yr)));
  return (int)(((double)d / 256.0) + (1.05 * e));
}

/* Median prediction with its distortion from MBs decided in this call. */
void orc_umhs_pred(const int16_t *mv, const int32_t *dist, int mbw, int col, int row, int *px, int *py,
                   int *pd) {
  int Ax = 0, Ay = 0, Ad = 0, Bx = 0, By = 0, Bd = 0, Cx = 0, Cy = 0, Cd = 0;
  const int hasB = row > 0, hasC = row > 0 && col + 1 < mbw;
  if (col > 0) { int k = row * mbw + col - 1; Ax = mv[2 * k]; Ay = mv[2 * k + 1]; Ad = dist[k]; }
  if (hasB) { int k = (row - 1) * mbw + col; Bx = mv[2 * k]; By = mv[2 * k + 1]; Bd = dist[k]; }
  if (hasC) {
    int k = (row - 1) * mbw + col + 1; Cx = mv[2 * k]; Cy = mv[2 * k + 1]; Cd = dist[k];
  } else if (col > 0 && row > 0) {
    int k = (row - 1) * mbw + col - 1; Cx = mv[2 * k]; Cy = mv[2 * k + 1]; Cd = dist[k];
  }
  if (!hasB && !hasC) { Bx = Cx = Ax; By = Cy = Ay; Bd = Cd = Ad; }
  const int x = median3(Ax, Bx, Cx), y = median3(Ay, By, Cy);
  int xd = Ad, yd = Ad;
  if ((x == Bx) && (Bd < xd)) xd = Bd;
  if ((x == Cx) && (Cd < xd)) xd = Cd;
  if ((y == By) && (Bd < yd)) yd = Bd;
  if ((y == Cx) && (Cd < yd)) yd = Cd; /* the reference compares with Cx here */
  *px = x;
  *py = y;
  *pd = yd < xd ? yd : xd;
}

typedef struct {
  const int16_t *cur;
  const plane_t *ref;
  int W, n, m;
  int xl, xr, yu, yd;
  int pX0, pY0; /* cost reference: the truncated full-pel predictor */
  int minD, minC, rmx, rmy;
} umhs_t;

/* One candidate at offset (j, i) from (mx, my) (the loops' common body). */
static void umhs_try(umhs_t *u, int mx, int my, int j, int i) {
  if ((i + my) < u->yu || (i + my) > u->yd || (j + mx) < u->xl || (j + mx) > u->xr) return;
  int blk[256];
  read_full(u->ref, u->n + mx + j, u->m + my + i, blk);
  const int d = blk_dist(u->cur, u->W, u->n, u->m, blk, 0);
  if (d <= u->minD) {
    const int c = umhs_cost(d, mx + j, my + i, u->pX0, u->pY0);
    if (c < u->minC) { u->minD = d; u->minC = c; u->rmx = j; u->rmy = i; }
  }
}

int orc_umhs_estimate(const int16_t *cur, const int16_t *refp, int W, int H, int qrange, int mode,
                      const int16_t *prev_mv, const uint8_t *prev_intra, const uint8_t *include,
                      int16_t *mv_out, int32_t *dist_out, long *avg_out) {
  if (W <= 0 || H <= 0 || (W & 15) || (H & 15) || qrange < 4) return -1;
  const int mbw = W / 16, mbh = H / 16, mRng = qrange / 4;
  plane_t ref = {refp, W, H};
  qwin_t *win = (qwin_t *)malloc(sizeof(qwin_t));
  int32_t *pdist = (int32_t *)calloc((size_t)mbw * mbh, sizeof(int32_t));
  if (!win || !pdist) { free(win); free(pdist); return -2; }
  memset(win, 0, sizeof(*win));
  int blk[256];
  long total = 0;
  int included = 0;
  for (int row = 0; row < mbh; row++)
    for (int col = 0; col < mbw; col++) {
      const int n = col * 16, m = row * 16, vec = row * mbw + col;
      int predX, predY, predD;
      orc_umhs_pred(mv_out, pdist, mbw, col, row, &predX, &predY, &predD);
      const int predXQ = predX % 4, predYQ = predY % 4;
      int predX0 = predX / 4, predY0 = predY / 4;
      const int oX0 = predX0, oY0 = predY0;
      if (predX0 + n > W) predX0 = 16;
      if (predX0 + n < -16) predX0 = -16;
      if (predY0 + m > H) predY0 = 16;
      if (predY0 + m < -16) predY0 = -16;
      const int trunc = (oX0 != predX0) || (oY0 != predY0);
      umhs_stat[7] += trunc;
      umhs_t u;
      u.cur = cur; u.ref = &ref; u.W = W; u.n = n; u.m = m;
      motion_range(n, m, W, H, mRng, 0, &u.xl, &u.xr, &u.yu, &u.yd);
      u.pX0 = predX0; u.pY0 = predY0;
      read_full(&ref, n + predX0, m + predY0, blk);
      int predVecD = blk_dist(cur, W, n, m, blk, 0);
      int mx = predX0, my = predY0;
      u.minD = predVecD;
      u.minC = predVecD / 256;
      u.rmx = u.rmy = 0;
      if (predX0 || predY0) {
        read_full(&ref, n, m, blk);
        const int zd = blk_dist(cur, W, n, m, blk, 0);
        const int zc = umhs_cost(zd, 0, 0, predX0, predY0);
        if (zc < u.minC) { u.minD = zd; u.minC = zc; mx = 0; my = 0; }
      }
      if (prev_mv && !prev_intra[vec]) {
        const int qX0 = prev_mv[2 * vec] / 4, qY0 = prev_mv[2 * vec + 1] / 4;
        if ((qX0 || qY0) && (qX0 != predX0) && (qY0 != predY0)) {
          read_full(&ref, n + qX0, m + qY0, blk);
          const int qd = blk_dist(cur, W, n, m, blk, 0);
          const int qc = umhs_cost(qd, qX0, qY0, predX0, predY0);
          if (qc < u.minC) { u.minD = qd; u.minC = qc; mx = qX0; my = qY0; }
          umhs_stat[6]++;
        }
      }
      int stage; /* 0: cross, 1: extended hexagon, 2: diamond */
      if (u.minD < 1000) stage = 2;
      else if ((u.minD < 4000) ||
               ((u.minD == predVecD) && (u.minD < (predD * 12 / 10)) && (u.minD > (predD * 8 / 10))))
        stage = 1;
      else stage = 0;
      umhs_stat[stage == 2 ? 9 : (stage == 1 ? 10 : 0)]++;
      if (stage == 0) {
        /* unsymmetrical cross (:582-612) */
        u.rmx = u.rmy = 0;
        int prior = u.minD;
        for (int w = 1; w < 16; w += 2)
          for (int x = 0; x < 4; x++) {
            const int i = w * U_CROSS[x][1], j = w * U_CROSS[x][0];
            if ((i < mRng / 2) && (i > -mRng / 2)) umhs_try(&u, mx, my, j, i);
          }
        mx += u.rmx; my += u.rmy;
        if (u.minD < 1000) stage = 2;
        else if ((u.minD < 2000) || ((u.minD < (prior * 9 / 10)) && (u.rmx || u.rmy) && (abs(u.rmx) <= 3) &&
                                     (abs(u.rmy) <= 3)))
          stage = 1;
        if (stage != 0) umhs_stat[11]++;
        if (stage == 0) {
          umhs_stat[1]++;
          /* 5x5 (:624-653); its winner moves the centre only through the
             hexagon stage below */
          const int sx = u.rmx, sy = u.rmy;
          u.rmx = u.rmy = 0;
          for (int w = 0; w < 24; w++) {
            const int i = U_SQ5[w][1], j = U_SQ5[w][0];
            const int done = (sx || sy) && (((sy == 0) && ((j == 2) || (j == -2))) ||
                                            ((sx == 0) && ((i == 2) || (i == -2))));
            if (!done) umhs_try(&u, mx, my, j, i);
          }
          if (u.minD < 1000) {
            stage = 2;
            umhs_stat[12]++;
          } else {
            umhs_stat[2]++;
            /* uneven multi-hexagon (:663-698) */
            prior = u.minD;
            for (int w = 1; w < 4; w++) {
              for (int x = 0; x < 16; x++) umhs_try(&u, mx, my, w * U_HEX16[x][0], w * U_HEX16[x][1]);
              if ((u.rmx || u.rmy) && ((u.minD < 2000) || (u.minD < (prior / 5)))) { umhs_stat[3]++; break; }
            }
            mx += u.rmx; my += u.rmy;
            stage = 1;
          }
        }
      }
      if (stage == 1) {
        /* extended hexagon until its centre wins (:706-737) */
        do {
          u.rmx = u.rmy = 0;
          for (int x = 0; x < 6; x++) umhs_try(&u, mx, my, U_EXTHEX[x][0], U_EXTHEX[x][1]);
          mx += u.rmx; my += u.rmy;
          umhs_stat[4] += (u.rmx || u.rmy);
        } while (u.rmx || u.rmy);
      }
      /* diamond until its centre wins (:740-770) */
      do {
        u.rmx = u.rmy = 0;
        for (int x = 0; x < 4; x++) umhs_try(&u, mx, my, U_CROSS[x][0], U_CROSS[x][1]);
        mx += u.rmx; my += u.rmy;
        umhs_stat[5] += (u.rmx || u.rmy);
      } while (u.rmx || u.rmy);

      /* sub-pel stages (:795-862), as the full-search estimator's */
      int minD = u.minD;
      int mvx = mx * 4, mvy = my * 4;
      if (mode != 2) {
        int hmx = 0, hmy = 0;
        load_half_window(win, &ref, n + mx, m + my);
        for (int k = 0; k < 8; k++) {
          const int ox = 2 * SUBPOS[k][0], oy = 2 * SUBPOS[k][1];
          window_read(win, ox, oy, blk);
          const int d = blk_dist(cur, W, n, m, blk, 0);
          if (d < minD) { minD = d; hmx = ox; hmy = oy; }
        }
        int qmx = hmx, qmy = hmy;
        if (mode != 1) {
          load_quarter_window(win, hmx, hmy);
          for (int k = 0; k < 8; k++) {
            const int ox = hmx + SUBPOS[k][0], oy = hmy + SUBPOS[k][1];
            window_read(win, ox, oy, blk);
            const int d = blk_dist(cur, W, n, m, blk, 0);
            if (d < minD) { minD = d; qmx = ox; qmy = oy; }
          }
        }
        mvx += qmx;
        mvy += qmy;
      }
      /* the predictor at quarter pel (:866-903) */
      const int rX = predX0 * 4 + predXQ, rY = predY0 * 4 + predYQ;
      if (predXQ || predYQ) {
        orc_quarter_read(refp, W, H, predX0 + n, predY0 + m, predXQ, predYQ, blk);
        predVecD = blk_dist(cur, W, n, m, blk, 0);
      } else if (trunc) {
        read_full(&ref, predX0 + n, predY0 + m, blk);
        predVecD = blk_dist(cur, W, n, m, blk, 0);
      }
      if (umhs_cost(predVecD, rX, rY, predX, predY) <= umhs_cost(minD, mvx, mvy, predX, predY)) {
        minD = predVecD; mvx = rX; mvy = rY;
        umhs_stat[8]++;
      }
      if (include && include[vec]) { included++; total += minD; }
      mv_out[2 * vec] = (int16_t)mvx;
      mv_out[2 * vec + 1] = (int16_t)mvy;
      pdist[vec] = minD;
      if (dist_out) dist_out[vec] = minD;
    }
  free(win);
  free(pdist);
  if (avg_out) *avg_out = included ? total / included : 0;
  return 0;
}

/* ------------------------------------------------------------------------
 * MotionEstimatorH264ImplCross::Estimate (src/CodecUtils/MotionEstimatorH264ImplCross.cpp:354-660):
 * a cross search whose arm halves from range/2 down to 1 around the running
 * best, the full-search estimator's sub-pel stages, and GetBestVector
 * (MotionEstimatorH264ImplCross.h:140-168) with lambda 0.9 over distortions
 * normalised by the path length.  Every distortion is the squared error
 * over the first 64 points of the estimator's sampling path
 * (Tsd16x16PartialPath, OverlayMem2Dv2.cpp:2495-2518); the search's
 * comparisons use Tsd16x16PartialPathLessThan (:2449-2493), which stops
 * after any batch of 16 points whose extrapolation to the whole path
 * exceeds the current minimum and returns that extrapolation.
 * ---------------------------------------------------------------------- */

/* The first 80 points of the sampling path (Cross.cpp:89-99; the FHS
 * estimator's unrolled path, FHS.cpp:1685-1790, is the same 80), as y*16+x. */
static const uint8_t CROSS_PATH[80] = {
    165, 93,  18,  237, 225, 25,  97,  159, 216, 46,  71,  138, 244, 160, 48,  5,
    188, 116, 12,  234, 195, 90,  223, 52,  135, 126, 214, 59,  146, 169, 31,  39,
    83,  248, 193, 104, 16,  156, 254, 79,  203, 102, 183, 112, 190, 7,   242, 124,
    50,  148, 44,  57,  220, 3,   229, 69,  208, 152, 178, 251, 26,  114, 111, 201,
    37,  197, 65,  91,  141, 121, 61,  133, 186, 246, 173, 72,  163, 29,  129, 14};

/* Squared error of each batch of 16 path points, block blk[256]. */
static void cross_batches(const int16_t *cur, int W, int n, int m, const int *blk, int b[4]) {
  for (int g = 0; g < 4; g++) {
    b[g] = 0;
    for (int k = 0; k < 16; k++) {
      const int p = CROSS_PATH[16 * g + k], y = p >> 4, x = p & 15;
      const int e = cur[(long)(m + y) * W + n + x] - blk[p];
      b[g] += e * e;
    }
  }
}

static int cross_full(const int b[4]) { return b[0] + b[1] + b[2] + b[3]; }

/* Branch counts since the last orc_cross_stats call: 0 a batch
 * extrapolation rejected a candidate whose whole-path sum would have been
 * accepted, 1 a tie taken by distance, 2 a tie kept, 3 the predictor
 * chosen, 4 the zero vector chosen, 5 a cross arm outside the window. */
static long cross_stat[6];
void orc_cross_stats(long *out) {
  for (int i = 0; i < 6; i++) { out[i] = cross_stat[i]; cross_stat[i] = 0; }
}

/* The LessThan variant's return value (4 batches of 16). */
static int cross_lt(const int b[4], int min) {
  int s = 0;
  for (int g = 0; g < 4; g++) {
    s += b[g];
    const int e = (s * 4) / (g + 1);
    if (g < 3 && e > min) return e;
  }
  return s;
}

int orc_cross_estimate(const int16_t *cur, const int16_t *refp, int W, int H, int qrange, int mode, int pred,
                       const uint8_t *include, int16_t *mv_out, int32_t *dist_out, long *avg_out) {
  if (W <= 0 || H <= 0 || (W & 15) || (H & 15) || qrange < 4) return -1;
  const int mbw = W / 16, mbh = H / 16, mRng = qrange / 4;
  plane_t ref = {refp, W, H};
  qwin_t *win = (qwin_t *)malloc(sizeof(qwin_t));
  if (!win) return -2;
  memset(win, 0, sizeof(*win));
  int blk[256], bt[4];
  long total = 0;
  int included = 0;
  for (int row = 0; row < mbh; row++)
    for (int col = 0; col < mbw; col++) {
      const int n = col * 16, m = row * 16, vec = row * mbw + col;
      int predX = 0, predY = 0;
      if (pred) orc_median_pred(mv_out, mbw, col, row, &predX, &predY);
      const int predXQ = predX % 4, predYQ = predY % 4;
      int predX0 = predX / 4, predY0 = predY / 4;
      const int pXr = predX < 0 ? (predX - 2) / 4 : (predX + 2) / 4;
      const int pYr = predY < 0 ? (predY - 2) / 4 : (predY + 2) / 4;
      read_full(&ref, n, m, blk);
      cross_batches(cur, W, n, m, blk, bt);
      const int zeroD = cross_full(bt);
      int minD = zeroD, mx = 0, my = 0;
      int xl, xr, yu, yd;
      motion_range(n, m, W, H, mRng, 0, &xl, &xr, &yu, &yd);
      for (int w = mRng / 2; w > 0; w /= 2) {
        int rmx = 0, rmy = 0;
        for (int x = 0; x < 4; x++) {
          const int i = w * U_CROSS[x][1], j = w * U_CROSS[x][0];
          if ((i + my) < yu || (i + my) > yd || (j + mx) < xl || (j + mx) > xr) { cross_stat[5]++; continue; }
          read_full(&ref, n + j + mx, m + i + my, blk);
          cross_batches(cur, W, n, m, blk, bt);
          const int d = cross_lt(bt, minD);
          if (d > minD && cross_full(bt) <= minD) cross_stat[0]++;
          if (d <= minD) {
            if (d == minD) {
              const int cx = mx + rmx - pXr, cy = my + rmy - pYr, nx = mx + j - pXr, ny = my + i - pYr;
              if ((cy * cy + cx * cx) - (ny * ny + nx * nx) < 0) { cross_stat[2]++; continue; }
              cross_stat[1]++;
            }
            minD = d; rmx = j; rmy = i;
          }
        }
        mx += rmx; my += rmy;
      }
      int mvx = mx * 4, mvy = my * 4;
      if (mode != 2) {
        int hmx = 0, hmy = 0;
        load_half_window(win, &ref, n + mx, m + my);
        for (int k = 0; k < 8; k++) {
          const int ox = 2 * SUBPOS[k][0], oy = 2 * SUBPOS[k][1];
          window_read(win, ox, oy, blk);
          cross_batches(cur, W, n, m, blk, bt);
          const int d = cross_lt(bt, minD);
          if (d > minD && cross_full(bt) < minD) cross_stat[0]++;
          if (d < minD) { minD = d; hmx = ox; hmy = oy; }
        }
        int qmx = hmx, qmy = hmy;
        if (mode != 1) {
          load_quarter_window(win, hmx, hmy);
          for (int k = 0; k < 8; k++) {
            const int ox = hmx + SUBPOS[k][0], oy = hmy + SUBPOS[k][1];
            window_read(win, ox, oy, blk);
            cross_batches(cur, W, n, m, blk, bt);
            const int d = cross_lt(bt, minD);
            if (d < minD) { minD = d; qmx = ox; qmy = oy; }
          }
        }
        mvx += qmx;
        mvy += qmy;
      }
      if (predX0 + n > W) predX0 = 16;
      if (predX0 + n < -16) predX0 = -16;
      if (predY0 + m > H) predY0 = 16;
      if (predY0 + m < -16) predY0 = -16;
      predX = predX0 * 4 + predXQ;
      predY = predY0 * 4 + predYQ;
      if (predXQ || predYQ)
        orc_quarter_read(refp, W, H, predX0 + n, predY0 + m, predXQ, predYQ, blk);
      else
        read_full(&ref, predX0 + n, predY0 + m, blk);
      cross_batches(cur, W, n, m, blk, bt);
      const int predD = cross_full(bt);
      /* GetBestVector over [best, pred, zero, zero, zero], lambda 0.9,
         norm 64 (the path length) */
      const int xs[5] = {mvx, predX, 0, 0, 0}, ys[5] = {mvy, predY, 0, 0, 0};
      const int ds[5] = {minD, predD, zeroD, zeroD, zeroD};
      int pos = 0;
      double cc = ((double)ds[0] / 64.0) +
                  (0.9 * sqrt((double)((xs[0] - predX) * (xs[0] - predX) + (ys[0] - predY) * (ys[0] - predY))));
      for (int k = 1; k < 5; k++) {
        const double nc = ((double)ds[k] / 64.0) +
                          (0.9 * sqrt((double)((xs[k] - predX) * (xs[k] - predX) + (ys[k] - predY) * (ys[k] - predY))));
        if (nc < cc) { cc = nc; pos = k; }
      }
      cross_stat[3] += pos == 1;
      cross_stat[4] += pos >= 2;
      if (include && include[vec]) { included++; total += ds[pos]; }
      mv_out[2 * vec] = (int16_t)xs[pos];
      mv_out[2 * vec + 1] = (int16_t)ys[pos];
      if (dist_out) dist_out[vec] = ds[pos];
    }
  free(win);
  if (avg_out) *avg_out = included ? total / included : 0;
  return 0;
}

/* ------------------------------------------------------------------------
 * MotionEstimatorH264ImplFHS::Estimate (src/CodecUtils/MotionEstimatorH264ImplFHS.cpp:497-905)
 * with HalfPelEstimate (:1109-1211), QuarterPelEstimate (:1213-1489), its
 * QuarterRead (:1491-1683, the H.264 quarter samples, as OverlayMem2Dv2's)
 * and Td16x16OptimalPathLessThan (:1685-1790): the path distortion over 80
 * points, returned as the extrapolation sum * 256 / points at the first of
 * the checkpoints 32, 40, ..., 72 where it exceeds the minimum, else at 80.
 * The integer cost is (d + 269 floor(sqrt(dist^2))) >> 8 (:222).
 *
 * The half-pel stage fills a cache of b / h / j samples that persists in
 * the estimator (and across Estimate calls) in five parts: h (rows -1..15,
 * cols 0..15), then, unless the minimum fell below 1000, h's side columns
 * and b (rows 0..15, cols -1..15), then, unless it fell below 1000 again,
 * b's side rows and j.  The quarter stage reads the cache whatever was
 * filled, so it can read samples an earlier MB left.  The cache is kept here
 * as the reference keeps it; entries never written are 0 here (the
 * reference's are uninitialised memory) and such reads are counted.
 * ---------------------------------------------------------------------- */
#define FHS_H 0
#define FHS_B 1
#define FHS_J 2

typedef struct {
  int c[3][18][18];
  unsigned char w[3][18][18]; /* written at least once */
} fhs_cache_t;

static long fhs_stat[8];
/* 0 uninitialised cache reads, 1 stale cache reads (written by an earlier
 * MB), 2 predictor early exit, 3 zero-vector early exit, 4 hexagon stage
 * run, 5 half-pel early exit after h, 6 after b, 7 quarter early exit. */
void orc_fhs_stats(long *out) {
  for (int i = 0; i < 8; i++) { out[i] = fhs_stat[i]; fhs_stat[i] = 0; }
}

static int fhs_path_lt(const int16_t *cur, int W, int n, int m, const int *blk, int min) {
  int dp = 0;
  for (int p = 0; p < 80; p++) {
    const int q = CROSS_PATH[p], e = cur[(long)(m + (q >> 4)) * W + n + (q & 15)] - blk[q];
    dp += e * e;
    const int k = p + 1;
    if (k >= 32 && k <= 72 && (k & 7) == 0) {
      const int pd = (dp * 256) / k;
      if (pd > min) return pd;
    }
  }
  return (dp * 256) / 80;
}

static int fhs_cost(int d, int x, int y, int xr, int yr) {
  const int mag = (int)sqrt((double)((x - xr) * (x - xr) + (y - yr) * (y - yr)));
  return (d + 269 * mag) >> 8;
}

/* The sample a quarter candidate reads from the cache; marks stale /
 * uninitialised reads.  fresh: this MB's fill generation of each part. */
typedef struct {
  fhs_cache_t *k;
  unsigned char mine[3][18][18]; /* written by the current MB */
} fhs_ctx_t;

static int fhs_rd(fhs_ctx_t *x, int t, int r, int c) {
  if (!x->k->w[t][r][c]) fhs_stat[0]++;
  else if (!x->mine[t][r][c]) fhs_stat[1]++;
  return x->k->c[t][r][c];
}
static void fhs_wr(fhs_ctx_t *x, int t, int r, int c, int v) {
  x->k->c[t][r][c] = v;
  x->k->w[t][r][c] = 1;
  x->mine[t][r][c] = 1;
}

/* HalfPelEstimate around full pel (x0, y0) (the MB at (n, m)). */
static int fhs_half(fhs_ctx_t *cx, const int16_t *cur, int W, const plane_t *ref, int n, int m, int x0, int y0,
                    int min, int *ohx, int *ohy) {
  int lmin = min, hx = 0, hy = 0;
  int dm = 0, dp = 0;
  for (int r = -1; r < 16; r++)
    for (int c = 0; c < 16; c++) {
      const int h = get_h(ref, x0 + c, y0 + r);
      fhs_wr(cx, FHS_H, r + 1, c + 1, h);
      if (dm <= lmin && r <= 14) { const int e = cur[(long)(m + r + 1) * W + n + c] - h; dm += e * e; }
      if (dp <= lmin && r >= 0) { const int e = cur[(long)(m + r) * W + n + c] - h; dp += e * e; }
    }
  if (dm < lmin) { lmin = dm; hx = 0; hy = -1; }
  if (dp < lmin) { lmin = dp; hx = 0; hy = 1; }
  if (lmin < 1000) { fhs_stat[5]++; goto done; }
  for (int r = -1; r < 16; r++) {
    fhs_wr(cx, FHS_H, r + 1, 0, get_h(ref, x0 - 1, y0 + r));
    fhs_wr(cx, FHS_H, r + 1, 17, get_h(ref, x0 + 16, y0 + r));
  }
  dm = dp = 0;
  for (int r = 0; r < 16; r++)
    for (int c = -1; c < 16; c++) {
      const int b = get_b(ref, x0 + c, y0 + r);
      fhs_wr(cx, FHS_B, r + 1, c + 1, b);
      if (dm <= lmin && c <= 14) { const int e = cur[(long)(m + r) * W + n + c + 1] - b; dm += e * e; }
      if (dp <= lmin && c >= 0) { const int e = cur[(long)(m + r) * W + n + c] - b; dp += e * e; }
    }
  if (dm < lmin) { lmin = dm; hx = -1; hy = 0; }
  if (dp < lmin) { lmin = dp; hx = 1; hy = 0; }
  if (lmin < 1000) { fhs_stat[6]++; goto done; }
  for (int c = -1; c < 16; c++) {
    fhs_wr(cx, FHS_B, 0, c + 1, get_b(ref, x0 + c, y0 - 1));
    fhs_wr(cx, FHS_B, 17, c + 1, get_b(ref, x0 + c, y0 + 16));
  }
  {
    int d[4] = {0, 0, 0, 0};
    for (int r = -1; r < 16; r++)
      for (int c = -1; c < 16; c++) {
        const int j = get_j(ref, x0 + c, y0 + r);
        fhs_wr(cx, FHS_J, r + 1, c + 1, j);
        if (d[0] <= lmin && r <= 14 && c <= 14) { const int e = cur[(long)(m + r + 1) * W + n + c + 1] - j; d[0] += e * e; }
        if (d[1] <= lmin && r <= 14 && c >= 0) { const int e = cur[(long)(m + r + 1) * W + n + c] - j; d[1] += e * e; }
        if (d[2] <= lmin && r >= 0 && c <= 14) { const int e = cur[(long)(m + r) * W + n + c + 1] - j; d[2] += e * e; }
        if (d[3] <= lmin && r >= 0 && c >= 0) { const int e = cur[(long)(m + r) * W + n + c] - j; d[3] += e * e; }
      }
    if (d[0] < lmin) { lmin = d[0]; hx = -1; hy = -1; }
    if (d[1] < lmin) { lmin = d[1]; hx = 1; hy = -1; }
    if (d[2] < lmin) { lmin = d[2]; hx = -1; hy = 1; }
    if (d[3] < lmin) { lmin = d[3]; hx = 1; hy = 1; }
  }
done:
  *ohx = hx;
  *ohy = hy;
  return lmin;
}

/* One quarter candidate's sample at block pel (r, c) (QuarterPelEstimate's
 * switch, :1258-1349): qx, qy in -3..3 around full pel (x0, y0). */
static int fhs_qsample(fhs_ctx_t *cx, const plane_t *ref, int x0, int y0, int qx, int qy, int r, int c) {
  const int mx = qx < 0 ? -1 : 0, my = qy < 0 ? -1 : 0;
  const int fx = qx - 4 * mx, fy = qy - 4 * my; /* 0..3 */
  const int cy = my + 1 + r, cc = mx + 1 + c;     /* cache index of the sample's own b/h/j */
  const int G = pix(ref, x0 + mx + c, y0 + my + r);
  switch (fy * 4 + fx) {
    case 1: return (G + fhs_rd(cx, FHS_B, cy, cc) + 1) >> 1;                                   /* a */
    case 3: return (pix(ref, x0 + mx + c + 1, y0 + my + r) + fhs_rd(cx, FHS_B, cy, cc) + 1) >> 1; /* c */
    case 4: return (G + fhs_rd(cx, FHS_H, cy, cc) + 1) >> 1;                                   /* d */
    case 12: return (pix(ref, x0 + mx + c, y0 + my + r + 1) + fhs_rd(cx, FHS_H, cy, cc) + 1) >> 1; /* n */
    case 5: return (fhs_rd(cx, FHS_B, cy, cc) + fhs_rd(cx, FHS_H, cy, cc) + 1) >> 1;           /* e */
    case 7: return (fhs_rd(cx, FHS_B, cy, cc) + fhs_rd(cx, FHS_H, cy, cc + 1) + 1) >> 1;       /* g */
    case 6: return (fhs_rd(cx, FHS_B, cy, cc) + fhs_rd(cx, FHS_J, cy, cc) + 1) >> 1;           /* f */
    case 14: return (fhs_rd(cx, FHS_B, cy + 1, cc) + fhs_rd(cx, FHS_J, cy, cc) + 1) >> 1;      /* q */
    case 9: return (fhs_rd(cx, FHS_H, cy, cc) + fhs_rd(cx, FHS_J, cy, cc) + 1) >> 1;           /* i */
    case 11: return (fhs_rd(cx, FHS_H, cy, cc + 1) + fhs_rd(cx, FHS_J, cy, cc) + 1) >> 1;      /* k */
    case 13: return (fhs_rd(cx, FHS_H, cy, cc) + fhs_rd(cx, FHS_B, cy + 1, cc) + 1) >> 1;      /* p */
    case 15: return (fhs_rd(cx, FHS_H, cy, cc + 1) + fhs_rd(cx, FHS_B, cy + 1, cc) + 1) >> 1;  /* r */
  }
  abort(); /* whole and half positions are not quarter candidates */
}

static int fhs_quarter(fhs_ctx_t *cx, const int16_t *cur, int W, const plane_t *ref, int n, int m, int x0, int y0,
                       int min, int *oqx, int *oqy) {
  int lmin = min, hx = 0, hy = 0, qx = 0, qy = 0;
  const int half = fhs_half(cx, cur, W, ref, n, m, x0, y0, min, &hx, &hy);
  if (half < min) {
    lmin = half;
    if (half < 1000) goto done;
  }
  for (int k = 0; k < 8; k++) {
    const int ox = 2 * hx + SUBPOS[k][0], oy = 2 * hy + SUBPOS[k][1];
    int q = 0;
    for (int r = 0; r < 16; r++)
      for (int c = 0; c < 16; c++) {
        const int e = cur[(long)(m + r) * W + n + c] - fhs_qsample(cx, ref, x0, y0, ox, oy, r, c);
        q += e * e;
      }
    if (q < lmin) {
      lmin = q; qx = ox; qy = oy;
      if (q < 1000) { fhs_stat[7]++; break; }
    }
  }
done:
  *oqx = qx;
  *oqy = qy;
  return lmin;
}

/* cache: 3*18*18 ints and 3*18*18 written flags carried from call to call
 * (both zero before the first; the reference allocates the cache in
 * Create()).  dist_out: the distortion the predictor keeps per MB. */
int orc_fhs_estimate(const int16_t *cur, const int16_t *refp, int W, int H, int qrange, int mode,
                     const uint8_t *include, int32_t *cache, uint8_t *cache_written, int16_t *mv_out,
                     int32_t *dist_out, long *avg_out) {
  if (W <= 0 || H <= 0 || (W & 15) || (H & 15) || qrange < 4) return -1;
  const int mbw = W / 16, mbh = H / 16, mRng = qrange / 4;
  plane_t ref = {refp, W, H};
  fhs_ctx_t *cx = (fhs_ctx_t *)calloc(1, sizeof(fhs_ctx_t));
  fhs_cache_t *kc = (fhs_cache_t *)calloc(1, sizeof(fhs_cache_t));
  int32_t *pdist = (int32_t *)calloc((size_t)mbw * mbh, sizeof(int32_t));
  if (!cx || !kc || !pdist) { free(cx); free(kc); free(pdist); return -2; }
  memcpy(kc->c, cache, sizeof(kc->c));
  memcpy(kc->w, cache_written, sizeof(kc->w));
  cx->k = kc;
  int blk[256];
  long total = 0;
  int included = 0;
  for (int row = 0; row < mbh; row++)
    for (int col = 0; col < mbw; col++) {
      const int n = col * 16, m = row * 16, vec = row * mbw + col;
      memset(cx->mine, 0, sizeof(cx->mine));
      int predX, predY, predD;
      orc_umhs_pred(mv_out, pdist, mbw, col, row, &predX, &predY, &predD);
      int pXQ = predX % 4, pYQ = predY % 4;
      int predX0 = predX / 4, predY0 = predY / 4;
      const int oX0 = predX0, oY0 = predY0;
      if (predX0 + n > W) predX0 = 16;
      if (predX0 + n < -16) predX0 = -16;
      if (predY0 + m > H) predY0 = 16;
      if (predY0 + m < -16) predY0 = -16;
      if (oX0 != predX0 || oY0 != predY0) { pXQ = 0; pYQ = 0; }
      const int rX = predX0 * 4 + pXQ, rY = predY0 * 4 + pYQ;
      int xl, xr, yu, yd;
      motion_range(n, m, W, H, mRng, 0, &xl, &xr, &yu, &yd);
      if (pXQ || pYQ) orc_quarter_read(refp, W, H, n + predX0, m + predY0, pXQ, pYQ, blk);
      else read_full(&ref, n + predX0, m + predY0, blk);
      const int predVecD = blk_dist(cur, W, n, m, blk, 0);
      int minD = predVecD, mx = predX0, my = predY0, minC = predVecD / 256;
      int mvx, mvy;
      if (predVecD < 1000) { mvx = rX; mvy = rY; fhs_stat[2]++; goto all_done; }
      if (rX || rY) {
        read_full(&ref, n, m, blk);
        const int zd = fhs_path_lt(cur, W, n, m, blk, minD);
        if (zd < minD) {
          const int zc = fhs_cost(zd, 0, 0, predX0, predY0);
          if (zc < minC) { minD = zd; minC = zc; mx = 0; my = 0; }
          if (zd < 1000) { mvx = 0; mvy = 0; fhs_stat[3]++; goto all_done; }
        }
      }
      {
        int stage; /* 0 hexagon, 1 extended hexagon, 2 diamond */
        if (minD < 1000) stage = 2;
        else if ((minD < 4000) || ((minD == predVecD) && (minD < (predD * 12 / 10)) && (minD > (predD * 8 / 10))))
          stage = 1;
        else stage = 0;
        int rmx = 0, rmy = 0;
#define FHS_TRY(J, I)                                                                                   \
  do {                                                                                                  \
    const int j_ = (J), i_ = (I);                                                                       \
    if ((i_ + my) >= yu && (i_ + my) <= yd && (j_ + mx) >= xl && (j_ + mx) <= xr) {                    \
      read_full(&ref, n + mx + j_, m + my + i_, blk);                                                   \
      const int d_ = fhs_path_lt(cur, W, n, m, blk, minD);                                              \
      if (d_ <= minD) {                                                                                 \
        const int c_ = fhs_cost(d_, mx + j_, my + i_, predX0, predY0);                                  \
        if (c_ < minC) { minD = d_; minC = c_; rmx = j_; rmy = i_; }                                    \
      }                                                                                                 \
    }                                                                                                   \
  } while (0)
        if (stage == 0) {
          fhs_stat[4]++;
          rmx = rmy = 0;
          for (int w = 1; w <= 4; w++)
            for (int x = 0; x < 16; x++) FHS_TRY(w * U_HEX16[x][0], w * U_HEX16[x][1]);
          mx += rmx; my += rmy;
        }
        if (stage <= 1) {
          rmx = rmy = 0;
          for (int x = 0; x < 6; x++) FHS_TRY(U_EXTHEX[x][0], U_EXTHEX[x][1]);
          mx += rmx; my += rmy;
        }
        rmx = rmy = 0;
        for (int x = 0; x < 4; x++) FHS_TRY(U_CROSS[x][0], U_CROSS[x][1]);
        mx += rmx; my += rmy;
#undef FHS_TRY
      }
      mvx = mx * 4;
      mvy = my * 4;
      if (mode == 0) {
        int qx = 0, qy = 0;
        const int nm = fhs_quarter(cx, cur, W, &ref, n, m, n + mx, m + my, minD, &qx, &qy);
        if (nm < minD) { minD = nm; mvx += qx; mvy += qy; }
      } else if (mode == 1) {
        int hx = 0, hy = 0;
        const int nm = fhs_half(cx, cur, W, &ref, n, m, n + mx, m + my, minD, &hx, &hy);
        if (nm < minD) { minD = nm; mvx += 2 * hx; mvy += 2 * hy; }
      }
      if (fhs_cost(predVecD, rX, rY, predX0, predY0) <= fhs_cost(minD, mvx, mvy, predX0, predY0)) {
        minD = predVecD; mvx = rX; mvy = rY;
      }
    all_done:
      if (include && include[vec]) { included++; total += minD; }
      mv_out[2 * vec] = (int16_t)mvx;
      mv_out[2 * vec + 1] = (int16_t)mvy;
      pdist[vec] = minD;
      if (dist_out) dist_out[vec] = minD;
    }
  memcpy(cache, kc->c, sizeof(kc->c));
  memcpy(cache_written, kc->w, sizeof(kc->w));
  free(cx);
  free(kc);
  free(pdist);
  if (avg_out) *avg_out = included ? total / included : 0;
  return 0;
}
