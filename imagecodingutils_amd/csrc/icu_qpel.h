// H.264 luma quarter-sample interpolation from an LDS patch, and the
// correctly rounded double-precision helpers of the final vector choice.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icu {

// Patch of 22 x 22 reference samples around a 16x16 block whose full-pel
// origin sits at patch (3, 3).
constexpr int kPatch = 22;

__device__ __forceinline__ int clip255(int x) { return x < 0 ? 0 : (x > 255 ? 255 : x); }

#define ICU_TAP6(a, b, c, d, e, f) ((a) - 5 * (b) + 20 * (c) + 20 * (d) - 5 * (e) + (f))

template <typename T>
__device__ __forceinline__ int vtap_p(const T* P, int x, int y) {
  return ICU_TAP6((int)P[(y - 2) * kPatch + x], (int)P[(y - 1) * kPatch + x], (int)P[y * kPatch + x],
                  (int)P[(y + 1) * kPatch + x], (int)P[(y + 2) * kPatch + x],
                  (int)P[(y + 3) * kPatch + x]);
}
template <typename T>
__device__ __forceinline__ int htap_p(const T* P, int x, int y) {
  const T* r = P + y * kPatch;
  return ICU_TAP6((int)r[x - 2], (int)r[x - 1], (int)r[x], (int)r[x + 1], (int)r[x + 2], (int)r[x + 3]);
}
template <typename T>
__device__ __forceinline__ int b_p(const T* P, int x, int y) { return clip255((htap_p(P, x, y) + 16) >> 5); }
template <typename T>
__device__ __forceinline__ int h_p(const T* P, int x, int y) { return clip255((vtap_p(P, x, y) + 16) >> 5); }
template <typename T>
__device__ __forceinline__ int j_p(const T* P, int x, int y) {
  int s = ICU_TAP6(vtap_p(P, x - 2, y), vtap_p(P, x - 1, y), vtap_p(P, x, y), vtap_p(P, x + 1, y),
                   vtap_p(P, x + 2, y), vtap_p(P, x + 3, y));
  return clip255((s + 512) >> 10);
}

// Quarter sample at patch full-pel (x, y) with fractions fx, fy in 0..3
// (OverlayMem2Dv2::QuarterRead, src/ImageUtils/OverlayMem2Dv2.cpp:729-1010;
// the 16 positions G, a..r of H.264 8.4.2.2.1).
template <typename T>
__device__ __forceinline__ int qpel_p(const T* P, int x, int y, int fx, int fy) {
  const int G = P[y * kPatch + x];
  switch (fx | (fy << 2)) {
    case 0: return G;
    case 2: return b_p(P, x, y);
    case 8: return h_p(P, x, y);
    case 10: return j_p(P, x, y);
    case 1: return (G + b_p(P, x, y) + 1) >> 1;
    case 3: return ((int)P[y * kPatch + x + 1] + b_p(P, x, y) + 1) >> 1;
    case 4: return (G + h_p(P, x, y) + 1) >> 1;
    case 12: return ((int)P[(y + 1) * kPatch + x] + h_p(P, x, y) + 1) >> 1;
    case 6: return (j_p(P, x, y) + b_p(P, x, y) + 1) >> 1;
    case 14: return (j_p(P, x, y) + b_p(P, x, y + 1) + 1) >> 1;
    case 9: return (j_p(P, x, y) + h_p(P, x, y) + 1) >> 1;
    case 11: return (j_p(P, x, y) + h_p(P, x + 1, y) + 1) >> 1;
    case 5: return (b_p(P, x, y) + h_p(P, x, y) + 1) >> 1;
    case 7: return (b_p(P, x, y) + h_p(P, x + 1, y) + 1) >> 1;
    case 13: return (h_p(P, x, y) + b_p(P, x, y + 1) + 1) >> 1;
    default: return (h_p(P, x + 1, y) + b_p(P, x, y + 1) + 1) >> 1;
  }
}

// Correctly rounded sqrt of a non-negative integer below 2^22: the hardware
// result is moved to whichever neighbour leaves the smaller exact residual.
__device__ __forceinline__ double sqrt_int_rn(int m) {
  const double x = (double)m;
  double y = __builtin_sqrt(x);
  // neighbouring doubles of a positive y: adjacent bit patterns
  const long long yb = __double_as_longlong(y);
  const double yl = __longlong_as_double(yb - 1), yh = __longlong_as_double(yb + 1);
  const double r = __builtin_fabs(__builtin_fma(-y, y, x));
  const double rl = __builtin_fabs(__builtin_fma(-yl, yl, x));
  const double rh = __builtin_fabs(__builtin_fma(-yh, yh, x));
  if (rl < r) y = yl;
  else if (rh < r) y = yh;
  return m == 0 ? 0.0 : y;
}

// Cost of GetBestVector (include/CodecUtils/MotionEstimatorH264ImplFull.h:
// 140-168): dist / 256 + 1.05 * sqrt(mag), each operation rounded on its own.
// dist / 256 of an integer is exact in a double, so it is the product with
// 2^-8 (one multiply instead of a correctly rounded division).
__device__ __forceinline__ double best_vector_cost(int dist, int mag) {
  const double a = __dmul_rn((double)dist, 0.00390625);
  const double b = __dmul_rn(1.05, sqrt_int_rn(mag));
  return __dadd_rn(a, b);
}

}  // namespace icu
