// The Cross fast estimator (icu_cross_*): MotionEstimatorH264ImplCross
// (src/CodecUtils/MotionEstimatorH264ImplCross.cpp:354-660) on the device.
//
// Per MB: a cross search whose arm halves from range/2 down to 1 around the
// running best, the full-search estimator's half / quarter stages and
// GetBestVector (MotionEstimatorH264ImplCross.h:140-168, lambda 0.9, norm =
// the path length).  Every distortion is the squared error over the first
// 64 points of the estimator's sampling path (Tsd16x16PartialPath,
// OverlayMem2Dv2.cpp:2495-2518); the comparisons use its LessThan form
// (:2449-2493), which stops after any batch of 16 points whose
// extrapolation to the whole path, (sum * 4) / batches done, exceeds the
// current minimum and returns that extrapolation.  Lane p of the wave holds
// path point p, so each quarter of the wave sums one batch.
//
// With the zero predictor MBs are independent (one wave per MB); with the
// H.264 median predictor MB rows run as a wavefront, as in the full-search
// estimator's decision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/icu_hip.h"
#include "icu_decide.h"
#include "icu_device.h"
#include "icu_fastwin.h"
#include "icu_internal.h"

using namespace icu;
using namespace icu::dec;

namespace {

// The first 64 points of the sampling path (Cross.cpp:89-99) as y*16 + x.
__constant__ uint8_t kPath[64] = {
    165, 93,  18,  237, 225, 25,  97,  159, 216, 46,  71,  138, 244, 160, 48,  5,
    188, 116, 12,  234, 195, 90,  223, 52,  135, 126, 214, 59,  146, 169, 31,  39,
    83,  248, 193, 104, 16,  156, 254, 79,  203, 102, 183, 112, 190, 7,   242, 124,
    50,  148, 44,  57,  220, 3,   229, 69,  208, 152, 178, 251, 26,  114, 111, 201};

constexpr int kSpin = 1 << 26;

struct CrossDesc {
  DecideDesc d;     // cur / ref planes, sizes, subpel depth
  int mrng;         // full-pel range (motionRange / 4)
  int median;       // 1: H.264 median predictor (wavefront), 0: zero
  int16_t* mv_out;
  int32_t* dist_out;
  int* flags;       // [0] row ticket, [1] timed out
};

// The four batch sums of one candidate from each lane's squared error.
struct Batches {
  int b[4];
  __device__ __forceinline__ int full() const { return b[0] + b[1] + b[2] + b[3]; }
  // Tsd16x16PartialPathLessThan's return value against min
  __device__ __forceinline__ int lt(int min) const {
    int s = 0;
#pragma unroll
    for (int g = 0; g < 4; g++) {
      s += b[g];
      if (g < 3 && (s * 4) / (g + 1) > min) return (s * 4) / (g + 1);
    }
    return s;
  }
};

// Reduce K candidates' per-lane squared errors to their batch sums.
// (each batch is one 16-lane row: a DPP row scan leaves its sum in the
// row's last lane)
template <int K>
__device__ __forceinline__ void batch_sums(int (&e)[K], Batches (&out)[K]) {
#pragma unroll
  for (int k = 0; k < K; k++) e[k] = row_scan(e[k]);
#pragma unroll
  for (int k = 0; k < K; k++)
#pragma unroll
    for (int g = 0; g < 4; g++) out[k].b[g] = __builtin_amdgcn_readlane(e[k], 16 * g + 15);
}

__device__ __forceinline__ int sq(int a, int b) { return (a - b) * (a - b); }

// GetBestVector's cost with lambda 0.9 and norm 64, correctly rounded
// (dist / 64 of an integer is exact: the product with 2^-6).
__device__ __forceinline__ double cross_cost(int dist, int mag) {
  return __dadd_rn(__dmul_rn((double)dist, 0.015625), __dmul_rn(0.9, sqrt_int_rn(mag)));
}

// The MB's search window, staged by stage_window before cross_mb.
__device__ __forceinline__ fw::Win stage_window(const CrossDesc& c, uint8_t* win, int mb, int lane) {
  const DecideDesc& d = c.d;
  const int n = (mb % d.mbw) * 16, m = (mb / d.mbw) * 16;
  const fw::Win w = fw::shape(n, m, d.W, d.H, c.mrng - 1);
  fw::issue(d, w, win, lane);
  return w;
}

// Returns the chosen vector as its output word (x low, y high), in every lane.
__device__ uint32_t cross_mb(const CrossDesc& c, WaveLds& s, const uint8_t* win, const fw::Win& w, int mb, int predX,
                             int predY, int lane) {
  const DecideDesc& d = c.d;
  const int col = mb % d.mbw, row = mb / d.mbw;
  const int n = col * 16, m = row * 16;
  const int pp = kPath[lane], py = pp >> 4, px = pp & 15;
  const int cv = d.cur[(long)(m + py) * d.cur_stride + n + px];
  fw::wait();
  const uint8_t* rp = d.ref + (long)(m + py) * d.ref_stride + n + px;  // this lane's pel at the zero vector
  // this lane's pel of the block at vector (x, y): from the window when staged
  auto at = [&](int x, int y) {
    return w.has(n + x, m + y, 16, 16, 0) ? fw::byte_at(win, w, n + x + px, m + y + py)
                                          : (int)rp[(long)y * d.ref_stride + x];
  };

  const int predXQ = predX % 4, predYQ = predY % 4;
  int predX0 = predX / 4, predY0 = predY / 4;
  const int pXr = predX < 0 ? (predX - 2) / 4 : (predX + 2) / 4;
  const int pYr = predY < 0 ? (predY - 2) / 4 : (predY + 2) / 4;

  int zeroD;
  {
    int e[1] = {sq(cv, at(0, 0))};
    Batches b[1];
    batch_sums(e, b);
    zeroD = b[0].full();
  }
  int minD = zeroD, mx = 0, my = 0;
  int xl, xr, yu, yd;
  motion_window(n, m, d.W, d.H, c.mrng - 1, &xl, &xr, &yu, &yd);
  constexpr int kCx[4] = {-1, 1, 0, 0}, kCy[4] = {0, 0, -1, 1};
  for (int arm = c.mrng / 2; arm > 0; arm /= 2) {
    bool ok[4];
    int e[4];
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const int i = arm * kCy[x], j = arm * kCx[x];
      ok[x] = !((i + my) < yu || (i + my) > yd || (j + mx) < xl || (j + mx) > xr);
      e[x] = ok[x] ? sq(cv, at(mx + j, my + i)) : 0;
    }
    Batches b[4];
    batch_sums(e, b);
    int rmx = 0, rmy = 0;
#pragma unroll
    for (int x = 0; x < 4; x++) {
      if (!ok[x]) continue;
      const int i = arm * kCy[x], j = arm * kCx[x];
      const int v = b[x].lt(minD);
      if (v <= minD) {
        if (v == minD) {
          const int cx = mx + rmx - pXr, cy = my + rmy - pYr, nx = mx + j - pXr, ny = my + i - pYr;
          if ((cy * cy + cx * cx) - (ny * ny + nx * nx) < 0) continue;
        }
        minD = v;
        rmx = j;
        rmy = i;
      }
    }
    mx += rmx;
    my += rmy;
  }

  int mvx = mx * 4, mvy = my * 4;
  if (d.subpel >= 1) {
    // the 8 half candidates and (quarter mode) the 8 quarter averages around
    // the winner, sampled at this lane's path point: its lattice samples
    // filtered in registers from the reference pels around it (the same
    // G / b / h / j as build_half_grid's lattice)
    int p[7][10], V[2][10];
    load_pels_at(d, n + mx + px, m + my + py, p, V);
    int hxh = 0, hyh = 0;
    {
      int e[8] = {sq(cv, lattice_at<-1, -1>(p, V, 0)), sq(cv, lattice_at<-1, 0>(p, V, 0)),
                  sq(cv, lattice_at<-1, 1>(p, V, 0)),  sq(cv, lattice_at<0, -1>(p, V, 0)),
                  sq(cv, lattice_at<0, 1>(p, V, 0)),   sq(cv, lattice_at<1, -1>(p, V, 0)),
                  sq(cv, lattice_at<1, 0>(p, V, 0)),   sq(cv, lattice_at<1, 1>(p, V, 0))};
      Batches b[8];
      batch_sums(e, b);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int v = b[k].lt(minD);
        if (v < minD) {
          minD = v;
          hxh = kSubX[k];
          hyh = kSubY[k];
        }
      }
    }
    int qmx = 2 * hxh, qmy = 2 * hyh;
    if (d.subpel >= 2) {
      int q[6];
      winner_samples_at(p, V, 0, hyh, hxh, q);
      const int uu = q[0], u0 = q[1], l0 = q[2], c0 = q[3], r0 = q[4], d0 = q[5];
      int e[8] = {sq(cv, (uu + c0 + 1) >> 1), sq(cv, (u0 + c0 + 1) >> 1), sq(cv, (u0 + r0 + 1) >> 1),
                  sq(cv, (l0 + c0 + 1) >> 1), sq(cv, (r0 + c0 + 1) >> 1), sq(cv, (l0 + d0 + 1) >> 1),
                  sq(cv, (d0 + c0 + 1) >> 1), sq(cv, (d0 + r0 + 1) >> 1)};
      Batches b[8];
      batch_sums(e, b);
      const int hx = 2 * hxh, hy = 2 * hyh;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int v = b[k].lt(minD);
        if (v < minD) {
          minD = v;
          qmx = hx + kSubX[k];
          qmy = hy + kSubY[k];
        }
      }
    }
    mvx += qmx;
    mvy += qmy;
  }

  // the predictor, truncated to one MB beyond the picture, at quarter pel
  if (predX0 + n > d.W) predX0 = 16;
  if (predX0 + n < -16) predX0 = -16;
  if (predY0 + m > d.H) predY0 = 16;
  if (predY0 + m < -16) predY0 = -16;
  const int pX = predX0 * 4 + predXQ, pY = predY0 * 4 + predYQ;
  int predD;
  {
    int e[1];
    if (predXQ || predYQ) {
      int fox = 0, fx = predXQ, foy = 0, fy = predYQ;
      if (fx < 0) { fox = -1; fx += 4; }
      if (fy < 0) { foy = -1; fy += 4; }
      // this lane's path pel's quarter sample, filtered in registers
      int q[4];
      qpel_row4(d.ref, d.ref_stride, n + predX0 + fox + px, m + predY0 + foy + py, fx, fy, q);
      e[0] = sq(cv, q[0]);
    } else {
      e[0] = sq(cv, at(predX0, predY0));
    }
    Batches b[1];
    batch_sums(e, b);
    predD = b[0].full();
    wave_sync();
  }
  // GetBestVector over [best, pred, zero, zero, zero]: strictly cheaper wins
  int outx = mvx, outy = mvy, outd = minD;
  double cc = cross_cost(minD, (mvx - pX) * (mvx - pX) + (mvy - pY) * (mvy - pY));
  const double cp = cross_cost(predD, 0);
  if (cp < cc) {
    cc = cp;
    outx = pX;
    outy = pY;
    outd = predD;
  }
  if (cross_cost(zeroD, pX * pX + pY * pY) < cc) {
    outx = outy = 0;
    outd = zeroD;
  }
  const uint32_t wv = (uint32_t)(uint16_t)(int16_t)outx | ((uint32_t)(uint16_t)(int16_t)outy << 16);
  if (lane == 0) {
    __hip_atomic_store((uint32_t*)(c.mv_out + 2 * mb), wv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    c.dist_out[mb] = outd;
  }
  return wv;
}

// Zero predictor: MBs are independent.
__global__ __launch_bounds__(64) void cross_parallel_kernel(CrossDesc c) {
  __shared__ WaveLds s;
  __shared__ __align__(16) uint8_t s_win[16];
  const int nmb = c.d.mbw * c.d.mbh;
  for (int mb = blockIdx.x; mb < nmb; mb += gridDim.x) {
    // no wait to hide the staging behind here: every read goes to the
    // (L2-resident) padded plane
    const fw::Win w{0, 0, 0, 0};
    cross_mb(c, s, s_win, w, mb, 0, 0, threadIdx.x);
    __syncthreads();
  }
}

// Median predictor: one wave per MB row, rows by ticket in start order,
// each waiting for the above-right MB of the row above.  The hand-off is the
// vector word itself: the host sets every word to kVecPending first, an MB
// polls its above and above-right (or above-left) words until both are
// final, and its own single vector store publishes it (one memory round
// trip per MB on the critical path; a progress counter took three).
__global__ __launch_bounds__(64) void cross_wavefront_kernel(CrossDesc c) {
  __shared__ WaveLds s;
  __shared__ __align__(16) uint8_t s_win[fw::kBytes];
  __shared__ int row_s;
  const DecideDesc& d = c.d;
  const int lane = threadIdx.x;
  if (lane == 0) row_s = atomicAdd(&c.flags[0], 1);
  __syncthreads();
  const int row = row_s;
  if (row >= d.mbh) return;
  const uint32_t* mvw = (const uint32_t*)c.mv_out;
  int Ax = 0, Ay = 0;  // the left neighbour: this wave's previous MB
  for (int col = 0; col < d.mbw; col++) {
    // the MB's window into LDS while it waits for its neighbours
    const fw::Win w = stage_window(c, s_win, row * d.mbw + col, lane);
    const int mb = row * d.mbw + col;
    int Bx = 0, By = 0, Cx = 0, Cy = 0;
    if (row > 0) {
      // lane 0: above; lane 1: above-right, or above-left in the last column
      const bool needC = col + 1 < d.mbw || col > 0;
      const int k = lane == 0 ? mb - d.mbw : (col + 1 < d.mbw ? mb - d.mbw + 1 : mb - d.mbw - 1);
      uint32_t wn = 0;
      if (lane == 0 || (lane == 1 && needC)) {
        int spins = 0;
        while ((wn = __hip_atomic_load(&mvw[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == kVecPending) {
          if (++spins > kSpin) {
            atomicExch(&c.flags[1], 1);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      const uint32_t wB = __builtin_amdgcn_readlane(wn, 0), wC = __builtin_amdgcn_readlane(wn, 1);
      Bx = (int)(int16_t)(wB & 0xFFFF);
      By = (int)(int16_t)(wB >> 16);
      if (needC) {
        Cx = (int)(int16_t)(wC & 0xFFFF);
        Cy = (int)(int16_t)(wC >> 16);
      }
    } else {
      Bx = Cx = Ax;
      By = Cy = Ay;
    }
    const uint32_t wv =
        __builtin_amdgcn_readfirstlane(cross_mb(c, s, s_win, w, mb, median3(Ax, Bx, Cx), median3(Ay, By, Cy), lane));
    Ax = (int)(int16_t)(wv & 0xFFFF);
    Ay = (int)(int16_t)(wv >> 16);
    __syncthreads();
  }
}

}  // namespace

struct icu_cross {
  int W = 0, H = 0, mbw = 0, mbh = 0, qrange = 0, mode = 0, median = 0;
  uint8_t* pad_base = nullptr;
  uint8_t* pad = nullptr;
  int stride = 0, pad_wgs = 0;
  int* d_wave = nullptr;  // row ticket and timeout flag
  uint8_t *d_cur = nullptr, *d_ref = nullptr;
  int16_t* d_mv = nullptr;
  int32_t* d_dist = nullptr;
  std::vector<uint8_t> h8;
  hipStream_t own = nullptr;
  ~icu_cross() {
    hipFree(pad_base);
    hipFree(d_wave);
    hipFree(d_cur);
    hipFree(d_ref);
    hipFree(d_mv);
    hipFree(d_dist);
    if (own) hipStreamDestroy(own);
  }
};

namespace {

#define CR_TRY(expr)                                                                                    \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess) return icu_fail(ICU_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

extern "C" {

int icu_cross_create(int width, int height, int motion_range, int mode, int predictor, icu_cross** out) {
  if (!out) return icu_fail(ICU_E_ARG, "null argument");
  if (width < 16 || height < 16 || width % 16 || height % 16)
    return icu_fail(ICU_E_ARG, "width and height must be positive multiples of 16");
  if (motion_range < 8 || motion_range > 4 * 64)
    return icu_fail(ICU_E_ARG, "motion range (quarter pels) must be 8..256");
  if (mode < 0 || mode > 2) return icu_fail(ICU_E_ARG, "mode must be 0 (quarter), 1 (half) or 2 (full pel)");
  if (predictor != 0 && predictor != 1) return icu_fail(ICU_E_ARG, "predictor must be 0 (zero) or 1 (median)");
  icu_cross* h = new icu_cross();
  h->W = width;
  h->H = height;
  h->mbw = width / 16;
  h->mbh = height / 16;
  h->qrange = motion_range;
  h->mode = mode;
  h->median = predictor;
  h->stride = ((width + 2 * kPad) + 255) & ~255;
  h->pad_wgs = pad_plane_workgroups(width, height);
  const size_t n = (size_t)h->mbw * h->mbh;
  if (hipMalloc(&h->pad_base, (size_t)h->stride * (height + 2 * kPad)) != hipSuccess ||
      hipMalloc(&h->d_wave, sizeof(int) * 2) != hipSuccess ||
      hipMalloc(&h->d_cur, (size_t)width * height) != hipSuccess ||
      hipMalloc(&h->d_ref, (size_t)width * height) != hipSuccess || hipMalloc(&h->d_mv, 4 * n) != hipSuccess ||
      hipMalloc(&h->d_dist, 4 * n) != hipSuccess || hipStreamCreate(&h->own) != hipSuccess) {
    delete h;
    return icu_fail(ICU_E_DEVICE, "device allocation failed");
  }
  h->pad = h->pad_base + (size_t)kPad * h->stride + kPad;
  *out = h;
  return ICU_OK;
}

int icu_cross_destroy(icu_cross* h) {
  delete h;
  return ICU_OK;
}

int icu_cross_estimate_device(icu_cross* h, const uint8_t* d_cur, const uint8_t* d_ref, int16_t* d_mv,
                              int32_t* d_dist, void* stream) {
  if (!h || !d_cur || !d_ref || !d_mv || !d_dist) return icu_fail(ICU_E_ARG, "null argument");
  if (((uintptr_t)d_mv & 3) || ((uintptr_t)d_dist & 3))
    return icu_fail(ICU_E_ARG, "vectors and distortions must be 4-byte aligned");
  const hipStream_t st = stream ? (hipStream_t)stream : h->own;
  PadSet ps{};
  ps.n = 1;
  ps.d[0] = PadDesc{d_ref, h->pad, h->W, h->H, h->W, h->stride, 0};
  launch_pad_set(ps, h->pad_wgs, st);
  CrossDesc c{};
  DecideDesc& d = c.d;
  d.cur = d_cur;
  d.ref = h->pad;
  d.cur_stride = h->W;
  d.ref_stride = h->stride;
  d.W = h->W;
  d.H = h->H;
  d.mbw = h->mbw;
  d.mbh = h->mbh;
  d.R = h->qrange / 4;
  d.L = d.R - 1;
  d.metric = kMetricSSD;
  d.subpel = h->mode == 0 ? 2 : (h->mode == 1 ? 1 : 0);
  c.mrng = h->qrange / 4;
  c.median = h->median;
  c.mv_out = d_mv;
  c.dist_out = d_dist;
  c.flags = h->d_wave;
  if (h->median) {
    launch_fill_i32((int*)d_mv, (int)dec::kVecPending, h->mbw * h->mbh, st);
    launch_fill_i32(h->d_wave, 0, 2, st);
    hipLaunchKernelGGL(cross_wavefront_kernel, dim3(h->mbh), dim3(64), 0, st, c);
  } else {
    const int nmb = h->mbw * h->mbh;
    hipLaunchKernelGGL(cross_parallel_kernel, dim3(std::min(nmb, 8192)), dim3(64), 0, st, c);
  }
  CR_TRY(hipGetLastError());
  return ICU_OK;
}

int icu_cross_check(icu_cross* h) {
  if (!h) return icu_fail(ICU_E_ARG, "null argument");
  int flags[2] = {0, 0};
  CR_TRY(hipDeviceSynchronize());
  if (h->median) CR_TRY(hipMemcpy(flags, h->d_wave, sizeof(flags), hipMemcpyDeviceToHost));
  return flags[1] ? icu_fail(ICU_E_STATE, "Cross wavefront timed out") : ICU_OK;
}

int icu_cross_estimate(icu_cross* h, const int16_t* cur, const int16_t* ref, const uint8_t* include, int16_t* mv,
                       int32_t* dist, long* avg) {
  if (!h || !cur || !ref || !mv) return icu_fail(ICU_E_ARG, "null argument");
  const size_t np = (size_t)h->W * h->H, nmb = (size_t)h->mbw * h->mbh;
  h->h8.resize(2 * np);
  for (size_t i = 0; i < np; i++) {
    if (cur[i] < 0 || cur[i] > 255 || ref[i] < 0 || ref[i] > 255)
      return icu_fail(ICU_E_RANGE, "input sample outside 0..255");
    h->h8[i] = (uint8_t)cur[i];
    h->h8[np + i] = (uint8_t)ref[i];
  }
  CR_TRY(hipMemcpyAsync(h->d_cur, h->h8.data(), np, hipMemcpyHostToDevice, h->own));
  CR_TRY(hipMemcpyAsync(h->d_ref, h->h8.data() + np, np, hipMemcpyHostToDevice, h->own));
  int rc = icu_cross_estimate_device(h, h->d_cur, h->d_ref, h->d_mv, h->d_dist, h->own);
  if (rc) return rc;
  std::vector<int32_t> dd(nmb);
  CR_TRY(hipMemcpyAsync(mv, h->d_mv, 4 * nmb, hipMemcpyDeviceToHost, h->own));
  CR_TRY(hipMemcpyAsync(dd.data(), h->d_dist, 4 * nmb, hipMemcpyDeviceToHost, h->own));
  // everything on the handle's own stream: objects driven from different
  // threads overlap on the device (a null-stream copy would wait for them all)
  int flags[2] = {0, 0};
  if (h->median)
    CR_TRY(hipMemcpyAsync(flags, h->d_wave, sizeof(flags), hipMemcpyDeviceToHost, h->own));
  CR_TRY(hipStreamSynchronize(h->own));
  if (flags[1]) return icu_fail(ICU_E_STATE, "Cross wavefront timed out");
  if (dist) std::copy(dd.begin(), dd.end(), dist);
  if (avg) {
    long long tot = 0;
    long cnt = 0;
    for (size_t i = 0; include && i < nmb; i++)
      if (include[i]) {
        tot += dd[i];
        cnt++;
      }
    *avg = cnt ? (long)(tot / cnt) : 0;
  }
  return ICU_OK;
}

}  // extern "C"
