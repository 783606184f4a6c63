/*
 * CPU oracle of the chroma half of MotionCompensatorH264ImplStd::Compensate
 * (src/CodecUtils/MotionCompensatorH264ImplStd.cpp:438-473) with
 * OverlayMem2Dv2::EighthRead (src/ImageUtils/OverlayMem2Dv2.cpp:1063-1114).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as the checker, never by the
 * product library.
 *
 * Per 16x16 MB with quarter-pel luma vector (mvx, mvy), the 8x8 chroma block
 * at (n/2, m/2) of each chroma plane is predicted from the reference chroma
 * plane at full-pel offset (mvx / 8, mvy / 8) with eighth fractions
 * (mvx % 8, mvy % 8), C truncation.  A negative fraction f reads one full pel
 * further left / above with fraction 8 + f.  Non-zero fractions interpolate
 * bilinearly: ((8-x)(8-y) A + x(8-y) B + (8-x)y C + x y D + 32) >> 6; zero
 * fractions copy.  Zero vectors leave the block as it is, which equals the
 * copy.  Reference samples outside the plane repeat its edge (the
 * reference's extended boundary).
 */
#include <stdlib.h>

#include "oracle.h"

static int cpx(const uint8_t *p, int W, int H, int x, int y) {
  x = x < 0 ? 0 : (x >= W ? W - 1 : x);
  y = y < 0 ? 0 : (y >= H ? H - 1 : y);
  return p[(long)y * W + x];
}

void orc_mc_chroma(const uint8_t *ref, int Wc, int Hc, const int16_t *mv, uint8_t *out) {
  const int mbw = Wc / 8, mbh = Hc / 8;
  for (int mr = 0; mr < mbh; mr++)
    for (int mc = 0; mc < mbw; mc++) {
      const int i = mr * mbw + mc;
      const int mvx = mv[2 * i], mvy = mv[2 * i + 1];
      int fx = mvx % 8, fy = mvy % 8;
      int ox = 8 * mc + mvx / 8, oy = 8 * mr + mvy / 8;
      if (fx < 0) { ox -= 1; fx += 8; }
      if (fy < 0) { oy -= 1; fy += 8; }
      for (int r = 0; r < 8; r++)
        for (int c = 0; c < 8; c++) {
          int v;
          if (fx == 0 && fy == 0) {
            v = cpx(ref, Wc, Hc, ox + c, oy + r);
          } else {
            const int A = cpx(ref, Wc, Hc, ox + c, oy + r), B = cpx(ref, Wc, Hc, ox + c + 1, oy + r);
            const int C = cpx(ref, Wc, Hc, ox + c, oy + r + 1), D = cpx(ref, Wc, Hc, ox + c + 1, oy + r + 1);
            v = ((8 - fx) * (8 - fy) * A + fx * (8 - fy) * B + (8 - fx) * fy * C + fx * fy * D + 32) >> 6;
          }
          out[(long)(8 * mr + r) * Wc + 8 * mc + c] = (uint8_t)v;
        }
    }
}
