// The UMHS fast estimator (icu_umhs_*): MotionEstimatorH264ImplUMHS
// (src/CodecUtils/MotionEstimatorH264ImplUMHS.cpp:446-940) on the device.
//
// The estimator is a per-MB walk whose next positions depend on the last
// stage's winner, and whose starting point is the median of the left, above
// and above-right MBs' vectors with their distortions
// (MacroBlockH264::GetMbMotionMedianPred, MacroBlockH264.cpp:818-915), so MB
// rows run as a wavefront: one wave per MB row, rows handed out by ticket in
// start order, each polling the row above's hand-off words (vector and
// distortion in one 64-bit word, kVecPending until the MB is final).
// Within an MB, every pattern stage evaluates all of its candidates at once
// (lane l holds the input block's row l / 4, columns 4 (l % 4) .. +3, and
// sums its 4 pixels of every candidate; the wave then reduces), and the
// reference's in-order running minimum is replayed over the sums.  The
// reference's partial sums (Tsd16x16LessThan returns early once the sum
// exceeds the current minimum) only ever reject candidates whose full sum
// exceeds it too, so full sums decide the same.  Sub-pel refinement and the
// predictor's quarter-pel read are the full-search estimator's.
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

// Pattern tables (UMHS.cpp:80-122), candidate order as the reference's.
__constant__ int8_t kSq5[24][2] = {{-1, 0}, {1, 0},  {0, -1},  {0, 1},  {-1, -1}, {1, -1}, {-1, 1}, {1, 1},
                                   {0, -2}, {-2, 0}, {2, 0},   {0, 2},  {-1, -2}, {1, -2}, {-2, -1}, {2, -1},
                                   {-2, 1}, {2, 1},  {-1, 2},  {1, 2},  {-2, -2}, {2, -2}, {-2, 2}, {2, 2}};
__constant__ int8_t kCross[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
__constant__ int8_t kHex16[16][2] = {{-4, 0}, {4, 0},  {-4, 1}, {4, 1},  {-4, -1}, {4, -1}, {-4, 2}, {4, 2},
                                     {-4, -2}, {4, -2}, {-2, 3}, {2, 3}, {-2, -3}, {2, -3}, {0, 4},  {0, -4}};
__constant__ int8_t kExtHex[6][2] = {{-2, 0}, {-1, -2}, {1, -2}, {2, 0}, {1, 2}, {-1, 2}};

constexpr int kSpin = 1 << 26;
// Each move of the hexagon / diamond loops lowers an integer cost that
// starts below 2^17 and never goes negative, so they end well before this.
constexpr int kMaxMoves = 1 << 18;

struct UmhsDesc {
  DecideDesc d;                // cur / ref planes, sizes, metric (SSD), subpel depth
  int mrng;                    // full-pel range (motionRange / 4)
  const int16_t* prev_mv;      // previous frame's co-located vectors (nullable)
  const uint8_t* prev_intra;   // and intra flags
  int16_t* mv_out;
  int32_t* dist_out;
  int* flags;                  // [0] row ticket, [1] timed out
  uint64_t* hand;              // per MB: vector word (low), distortion (high)
};

// MEH264IUMHS_COST: distortion / 256 + 1.05 x distance, truncated.
__device__ __forceinline__ int umhs_cost(int dist, int dx, int dy) {
  return (int)best_vector_cost(dist, dx * dx + dy * dy);
}

// The MB's staged search window (LDS) and its shape.
struct WinRef {
  const uint8_t* lds;
  fw::Win w;
};

// This lane's SSD over its 4 pixels for the block at full pel (x, y), read
// from the staged window when it lies inside.
__device__ __forceinline__ int ssd4(const DecideDesc& d, const WinRef& wr, int x, int y, uint32_t cw, int lane) {
  const uint32_t rw = fw::row4(d, wr.lds, wr.w, wr.w.has(x, y, 16, 16, 0), x, y, lane);
  int acc = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int e = (int)((cw >> (8 * i)) & 0xFFu) - (int)((rw >> (8 * i)) & 0xFFu);
    acc += e * e;
  }
  return acc;
}

__device__ __forceinline__ int block_ssd(const DecideDesc& d, const WinRef& wr, int x, int y, uint32_t cw, int lane) {
  return wave_sum(ssd4(d, wr, x, y, cw, lane));
}

// The running minimum of the search (minDiff, minCost, offset (rmx, rmy)
// from the stage centre), wave-uniform.
struct Best {
  int minD, minC, rmx, rmy;
};

// Up to 8 candidates at offsets (ox[k], oy[k]) from (mx, my), those with
// ok[k] inside the window, summed at once, then taken in order as the
// reference's loop body (UMHS.cpp:598-609): a candidate no worse in
// distortion and strictly cheaper in cost replaces the best.
__device__ __forceinline__ void try8(const DecideDesc& d, const WinRef& wr, int n, int m, int mx, int my,
                                     const int (&ox)[8], const int (&oy)[8], const bool (&ok)[8], int cnt, int pX0,
                                     int pY0, uint32_t cw, int lane, Best& b) {
  int acc[8];
#pragma unroll
  for (int k = 0; k < 8; k++)
    acc[k] = (k < cnt && ok[k]) ? ssd4(d, wr, n + mx + ox[k], m + my + oy[k], cw, lane) : 0;
  wave_sum8(acc);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (k < cnt && ok[k] && acc[k] <= b.minD) {
      const int c = umhs_cost(acc[k], mx + ox[k] - pX0, my + oy[k] - pY0);
      if (c < b.minC) {
        b.minD = acc[k];
        b.minC = c;
        b.rmx = ox[k];
        b.rmy = oy[k];
      }
    }
  }
}

// A stage's candidates one per lane (lane k the stage's k-th, ok: inside
// the window): each lane sums its candidate's SSD over the whole block from
// the staged window, as sum c^2 + sum r^2 - 2 sum c r with v_dot4_u32_u8
// (the block rows read from s.cur4, the same address in every lane), then
// the reference's loop body (UMHS.cpp:598-609) is replayed in order by
// ballot: the first candidate after the last accepted one that is no worse
// in distortion and strictly cheaper in cost replaces the best, until none
// is left.  One pass for a stage of up to 64 candidates, where try8 takes
// one per 8.  csq: the block's sum of squares.  Lanes without a candidate
// read the zero vector's block (always staged) and take no part.
__device__ __forceinline__ void try_lanes(const WinRef& wr, const WaveLds& s, int n, int m, int mx, int my, int ox,
                                          int oy, bool ok, uint32_t csq, int pX0, int pY0, int lane, Best& b) {
  const int x = ok ? n + mx + ox : n, y = ok ? m + my + oy : m;
  const int xs = x - wr.w.x0, sh = xs & 3;
  const uint8_t* rowp = wr.lds + (y - wr.w.y0) * wr.w.cols + (xs - sh);
  uint32_t cr = 0, rr = 0;
#pragma unroll 4
  for (int r = 0; r < 16; r++) {
    const uint32_t* q = (const uint32_t*)(rowp + r * wr.w.cols);
    uint32_t v[5];
#pragma unroll
    for (int j = 0; j < 5; j++) v[j] = q[j];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t a = __builtin_amdgcn_alignbyte(v[k + 1], v[k], sh);
      cr = __builtin_amdgcn_udot4(s.cur4[4 * r + k], a, cr, false);
      rr = __builtin_amdgcn_udot4(a, a, rr, false);
    }
  }
  const int acc = (int)(csq + rr - 2u * cr);
  const int cst = ok ? umhs_cost(acc, mx + ox - pX0, my + oy - pY0) : 0x7FFFFFFF;
  int start = 0;
  for (;;) {
    const uint64_t take = __ballot(ok && lane >= start && acc <= b.minD && cst < b.minC);
    if (take == 0) break;
    const int k = __builtin_ctzll(take);
    b.minD = __builtin_amdgcn_readlane(acc, k);
    b.minC = __builtin_amdgcn_readlane(cst, k);
    b.rmx = __builtin_amdgcn_readlane(ox, k);
    b.rmy = __builtin_amdgcn_readlane(oy, k);
    start = k + 1;
  }
}

struct Window {
  int xl, xr, yu, yd;
  __device__ __forceinline__ bool in(int x, int y) const { return y >= yu && y <= yd && x >= xl && x <= xr; }
};

// A table of cnt offsets (a multiple of 8 or fewer than 8) scaled by w,
// in groups of 8.
__device__ __forceinline__ void try_table(const DecideDesc& d, const WinRef& wr, const int8_t (*tab)[2], int cnt, int w,
                                          int n, int m, int mx, int my, const Window& win, int pX0, int pY0,
                                          uint32_t cw, int lane, Best& b) {
  for (int g = 0; g < cnt; g += 8) {
    int ox[8], oy[8];
    bool ok[8];
    const int c = min(8, cnt - g);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int t = k < c ? g + k : g;
      ox[k] = w * tab[t][0];
      oy[k] = w * tab[t][1];
      ok[k] = k < c && win.in(mx + ox[k], my + oy[k]);
    }
    try8(d, wr, n, m, mx, my, ox, oy, ok, c, pX0, pY0, cw, lane, b);
  }
}

__device__ __forceinline__ int median3i(int x, int y, int z) { return median3(x, y, z); }

__global__ __launch_bounds__(64) void umhs_kernel(UmhsDesc u) {
  __shared__ WaveLds s;
  __shared__ __align__(16) uint8_t s_win[fw::kBytes];
  __shared__ int row_s;
  const DecideDesc& d = u.d;
  const int lane = threadIdx.x;
  if (lane == 0) row_s = atomicAdd(&u.flags[0], 1);
  __syncthreads();
  const int row = row_s;
  if (row >= d.mbh) return;
  int Ax = 0, Ay = 0, Ad = 0;  // the left neighbour: this wave's previous MB
  for (int col = 0; col < d.mbw; col++) {
    const int n = col * 16, m = row * 16;
    // the MB's search window into LDS and its block into registers, while
    // the MB waits for its neighbours
    WinRef wr{s_win, fw::shape(n, m, d.W, d.H, u.mrng - 1)};
    fw::issue(d, wr.w, s_win, lane);
    const uint32_t cw = *(const uint32_t*)(d.cur + (long)(m + (lane >> 2)) * d.cur_stride + n + 4 * (lane & 3));
    const int mb = row * d.mbw + col;
    // lane 0: above; lane 1: above-right, or above-left in the last column;
    // one 64-bit word each, polled until its vector is no longer pending
    uint64_t hw = 0;
    const bool needC = row > 0 && (col + 1 < d.mbw || col > 0);
    if (row > 0 && (lane == 0 || (lane == 1 && needC))) {
      const int k = lane == 0 ? mb - d.mbw : (col + 1 < d.mbw ? mb - d.mbw + 1 : mb - d.mbw - 1);
      int spins = 0;
      while ((uint32_t)(hw = __hip_atomic_load(&u.hand[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ==
             kVecPending) {
        if (++spins > kSpin) {
          atomicExch(&u.flags[1], 1);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    fw::wait();
    // ---- predictor with distortion (MacroBlockH264.cpp:818-915)
    int Bx = 0, By = 0, Bd = 0, Cx = 0, Cy = 0, Cd = 0;
    if (row > 0) {
      const uint32_t bl = __builtin_amdgcn_readlane((uint32_t)hw, 0), bh = __builtin_amdgcn_readlane((uint32_t)(hw >> 32), 0);
      const uint32_t cl = __builtin_amdgcn_readlane((uint32_t)hw, 1), ch = __builtin_amdgcn_readlane((uint32_t)(hw >> 32), 1);
      Bx = (int)(int16_t)(bl & 0xFFFF);
      By = (int)(int16_t)(bl >> 16);
      Bd = (int)bh;
      if (needC) {
        Cx = (int)(int16_t)(cl & 0xFFFF);
        Cy = (int)(int16_t)(cl >> 16);
        Cd = (int)ch;
      }
    } else {
      Bx = Cx = Ax;
      By = Cy = Ay;
      Bd = Cd = Ad;
    }
    const int predX = median3i(Ax, Bx, Cx), predY = median3i(Ay, By, Cy);
    int xd = Ad, yd = Ad;
    if (predX == Bx && Bd < xd) xd = Bd;
    if (predX == Cx && Cd < xd) xd = Cd;
    if (predY == By && Bd < yd) yd = Bd;
    if (predY == Cx && Cd < yd) yd = Cd;  // the reference compares the y median with C's x
    const int predD = yd < xd ? yd : xd;

    // ---- starting points (UMHS.cpp:493-565)
    s.cur4[lane] = cw;
    const uint32_t csq = (uint32_t)wave_sum((int)__builtin_amdgcn_udot4(cw, cw, 0u, false));
    wave_sync();  // s.cur4 is read across lanes by try_lanes
    const int predXQ = predX % 4, predYQ = predY % 4;
    int predX0 = predX / 4, predY0 = predY / 4;
    const int oX0 = predX0, oY0 = predY0;
    if (predX0 + n > d.W) predX0 = 16;
    if (predX0 + n < -16) predX0 = -16;
    if (predY0 + m > d.H) predY0 = 16;
    if (predY0 + m < -16) predY0 = -16;
    const bool trunc = oX0 != predX0 || oY0 != predY0;
    Window win;
    motion_window(n, m, d.W, d.H, u.mrng - 1, &win.xl, &win.xr, &win.yu, &win.yd);
    int predVecD = block_ssd(d, wr, n + predX0, m + predY0, cw, lane);
    int mx = predX0, my = predY0;
    Best b{predVecD, predVecD / 256, 0, 0};
    if (predX0 || predY0) {
      const int zd = block_ssd(d, wr, n, m, cw, lane);
      const int zc = umhs_cost(zd, -predX0, -predY0);
      if (zc < b.minC) {
        b.minD = zd;
        b.minC = zc;
        mx = my = 0;
      }
    }
    if (u.prev_mv && !u.prev_intra[mb]) {
      const uint32_t w = *(const uint32_t*)(u.prev_mv + 2 * mb);
      // a previous vector must keep the block within the reference's 16-pel
      // border (the reference reads outside its buffer otherwise); here one
      // beyond it is brought to the border, so every later stage centred on
      // it stays inside the padded plane
      const int qX0 = min(max((int)(int16_t)(w & 0xFFFF) / 4, -(n + 16)), d.W - n);
      const int qY0 = min(max((int)(int16_t)(w >> 16) / 4, -(m + 16)), d.H - m);
      if ((qX0 || qY0) && qX0 != predX0 && qY0 != predY0) {
        const int qd = block_ssd(d, wr, n + qX0, m + qY0, cw, lane);
        const int qc = umhs_cost(qd, qX0 - predX0, qY0 - predY0);
        if (qc < b.minC) {
          b.minD = qd;
          b.minC = qc;
          mx = qX0;
          my = qY0;
        }
      }
    }

    // ---- full-pel stages (UMHS.cpp:571-770)
    int stage;  // 0 cross, 1 extended hexagon, 2 diamond
    if (b.minD < 1000) stage = 2;
    else if (b.minD < 4000 || (b.minD == predVecD && b.minD < predD * 12 / 10 && b.minD > predD * 8 / 10)) stage = 1;
    else stage = 0;
    if (stage == 0) {
      // unsymmetrical cross: horizontal arms to 15, vertical within
      // +-(range / 2) exclusive
      b.rmx = b.rmy = 0;
      int prior = b.minD;
      const int vlim = u.mrng / 2;
      {
        // 32 candidates: w = 1..15 odd, the four directions of each w in turn
        const int k = lane & 31, w = 2 * (k >> 2) + 1;
        const int ox = w * kCross[k & 3][0], oy = w * kCross[k & 3][1];
        const bool ok = lane < 32 && oy < vlim && oy > -vlim && win.in(mx + ox, my + oy);
        try_lanes(wr, s, n, m, mx, my, ox, oy, ok, csq, predX0, predY0, lane, b);
      }
      mx += b.rmx;
      my += b.rmy;
      if (b.minD < 1000) stage = 2;
      else if (b.minD < 2000 ||
               (b.minD < prior * 9 / 10 && (b.rmx || b.rmy) && abs(b.rmx) <= 3 && abs(b.rmy) <= 3))
        stage = 1;
      if (stage == 0) {
        // 5x5 without the cross's tested +-2 arm; its winner moves the
        // centre only through the hexagon stage
        const int sx = b.rmx, sy = b.rmy;
        b.rmx = b.rmy = 0;
        {
          const int k = lane < 24 ? lane : 0;
          const int ox = kSq5[k][0], oy = kSq5[k][1];
          const bool done = (sx || sy) && ((sy == 0 && (ox == 2 || ox == -2)) || (sx == 0 && (oy == 2 || oy == -2)));
          const bool ok = lane < 24 && !done && win.in(mx + ox, my + oy);
          try_lanes(wr, s, n, m, mx, my, ox, oy, ok, csq, predX0, predY0, lane, b);
        }
        if (b.minD < 1000) {
          stage = 2;
        } else {
          // uneven multi-hexagon, scales 1..3, leaving after a scale that
          // improved enough
          prior = b.minD;
          for (int w = 1; w < 4; w++) {
            const int k = lane & 15;
            const int ox = w * kHex16[k][0], oy = w * kHex16[k][1];
            const bool ok = lane < 16 && win.in(mx + ox, my + oy);
            try_lanes(wr, s, n, m, mx, my, ox, oy, ok, csq, predX0, predY0, lane, b);
            if ((b.rmx || b.rmy) && (b.minD < 2000 || b.minD < prior / 5)) break;
          }
          mx += b.rmx;
          my += b.rmy;
          stage = 1;
        }
      }
    }
    if (stage == 1) {
      for (int it = 0; it < kMaxMoves; it++) {
        b.rmx = b.rmy = 0;
        try_table(d, wr, kExtHex, 6, 1, n, m, mx, my, win, predX0, predY0, cw, lane, b);
        mx += b.rmx;
        my += b.rmy;
        if (!(b.rmx || b.rmy)) break;
      }
    }
    for (int it = 0; it < kMaxMoves; it++) {
      b.rmx = b.rmy = 0;
      try_table(d, wr, kCross, 4, 1, n, m, mx, my, win, predX0, predY0, cw, lane, b);
      mx += b.rmx;
      my += b.rmy;
      if (!(b.rmx || b.rmy)) break;
    }

    // ---- sub-pel stages and the predictor at quarter pel (UMHS.cpp:795-903)
    int minD = b.minD, mvx = mx * 4, mvy = my * 4;
    if (d.subpel >= 1) refine_subpel(d, s, n, m, mx, my, cw, lane, minD, mvx, mvy);
    const int rX = predX0 * 4 + predXQ, rY = predY0 * 4 + predYQ;
    if (predXQ || predYQ) predVecD = quarter_dist(d, s, n + predX0, m + predY0, predXQ, predYQ, lane);
    else if (trunc) predVecD = block_ssd(d, wr, n + predX0, m + predY0, cw, lane);
    if (umhs_cost(predVecD, rX - predX, rY - predY) <= umhs_cost(minD, mvx - predX, mvy - predY)) {
      minD = predVecD;
      mvx = rX;
      mvy = rY;
    }
    if (lane == 0) {
      const uint32_t w = (uint32_t)(uint16_t)(int16_t)mvx | ((uint32_t)(uint16_t)(int16_t)mvy << 16);
      *(uint32_t*)(u.mv_out + 2 * mb) = w;
      u.dist_out[mb] = minD;
      __hip_atomic_store(&u.hand[mb], (uint64_t)w | ((uint64_t)(uint32_t)minD << 32), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    Ax = __builtin_amdgcn_readfirstlane(mvx);
    Ay = __builtin_amdgcn_readfirstlane(mvy);
    Ad = __builtin_amdgcn_readfirstlane(minD);
    __syncthreads();
  }
}

}  // namespace

struct icu_umhs {
  int W = 0, H = 0, mbw = 0, mbh = 0, qrange = 0, mode = 0;
  uint8_t* pad_base = nullptr;
  uint8_t* pad = nullptr;
  int stride = 0, pad_wgs = 0;
  int* d_wave = nullptr;  // ticket and timeout flag, then one 64-bit hand-off word per MB
  // staging for the host entry point
  uint8_t *d_cur = nullptr, *d_ref = nullptr, *d_prev_intra = nullptr;
  int16_t *d_prev = nullptr, *d_mv = nullptr;
  int32_t* d_dist = nullptr;
  std::vector<uint8_t> h8;
  hipStream_t own = nullptr;
  ~icu_umhs() {
    hipFree(pad_base);
    hipFree(d_wave);
    hipFree(d_cur);
    hipFree(d_ref);
    hipFree(d_prev_intra);
    hipFree(d_prev);
    hipFree(d_mv);
    hipFree(d_dist);
    if (own) hipStreamDestroy(own);
  }
};

namespace {

#define UM_TRY(expr)                                                                                    \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess) return icu_fail(ICU_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

extern "C" {

int icu_umhs_create(int width, int height, int motion_range, int mode, icu_umhs** out) {
  if (!out) return icu_fail(ICU_E_ARG, "null argument");
  if (width < 16 || height < 16 || width % 16 || height % 16)
    return icu_fail(ICU_E_ARG, "width and height must be positive multiples of 16");
  if (motion_range < 8 || motion_range > 4 * 64)
    return icu_fail(ICU_E_ARG, "motion range (quarter pels) must be 8..256");
  if (mode < 0 || mode > 2) return icu_fail(ICU_E_ARG, "mode must be 0 (quarter), 1 (half) or 2 (full pel)");
  icu_umhs* h = new icu_umhs();
  h->W = width;
  h->H = height;
  h->mbw = width / 16;
  h->mbh = height / 16;
  h->qrange = motion_range;
  h->mode = mode;
  h->stride = ((width + 2 * kPad) + 255) & ~255;
  h->pad_wgs = pad_plane_workgroups(width, height);
  const size_t n = (size_t)h->mbw * h->mbh;
  if (hipMalloc(&h->pad_base, (size_t)h->stride * (height + 2 * kPad)) != hipSuccess ||
      hipMalloc(&h->d_wave, sizeof(int) * (2 + 2 * n)) != hipSuccess ||
      hipMalloc(&h->d_cur, (size_t)width * height) != hipSuccess ||
      hipMalloc(&h->d_ref, (size_t)width * height) != hipSuccess ||
      hipMalloc(&h->d_prev_intra, n) != hipSuccess || hipMalloc(&h->d_prev, 4 * n) != hipSuccess ||
      hipMalloc(&h->d_mv, 4 * n) != hipSuccess || hipMalloc(&h->d_dist, 4 * n) != hipSuccess ||
      hipStreamCreate(&h->own) != hipSuccess) {
    delete h;
    return icu_fail(ICU_E_DEVICE, "device allocation failed");
  }
  h->pad = h->pad_base + (size_t)kPad * h->stride + kPad;
  *out = h;
  return ICU_OK;
}

int icu_umhs_destroy(icu_umhs* h) {
  delete h;
  return ICU_OK;
}

int icu_umhs_estimate_device(icu_umhs* h, const uint8_t* d_cur, const uint8_t* d_ref, const int16_t* d_prev_mv,
                             const uint8_t* d_prev_intra, int16_t* d_mv, int32_t* d_dist, void* stream) {
  if (!h || !d_cur || !d_ref || !d_mv || !d_dist) return icu_fail(ICU_E_ARG, "null argument");
  if ((d_prev_mv == nullptr) != (d_prev_intra == nullptr))
    return icu_fail(ICU_E_ARG, "previous vectors and intra flags go together");
  if (((uintptr_t)d_cur & 15) || ((uintptr_t)d_mv & 3) || ((uintptr_t)d_dist & 3) || ((uintptr_t)d_prev_mv & 3))
    return icu_fail(ICU_E_ARG, "d_cur must be 16-byte aligned, vectors and distortions 4-byte aligned");
  const hipStream_t st = stream ? (hipStream_t)stream : h->own;
  PadSet ps{};
  ps.n = 1;
  ps.d[0] = PadDesc{d_ref, h->pad, h->W, h->H, h->W, h->stride, 0};
  launch_pad_set(ps, h->pad_wgs, st);
  launch_fill_i32(h->d_wave, 0, 2, st);
  launch_fill_i32(h->d_wave + 2, (int)dec::kVecPending, 2 * h->mbw * h->mbh, st);
  UmhsDesc u{};
  DecideDesc& d = u.d;
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
  u.mrng = h->qrange / 4;
  u.prev_mv = d_prev_mv;
  u.prev_intra = d_prev_intra;
  u.mv_out = d_mv;
  u.dist_out = d_dist;
  u.flags = h->d_wave;
  u.hand = (uint64_t*)(h->d_wave + 2);
  hipLaunchKernelGGL(umhs_kernel, dim3(h->mbh), dim3(64), 0, st, u);
  UM_TRY(hipGetLastError());
  return ICU_OK;
}

int icu_umhs_check(icu_umhs* h) {
  if (!h) return icu_fail(ICU_E_ARG, "null argument");
  int flags[2] = {0, 0};
  UM_TRY(hipDeviceSynchronize());
  UM_TRY(hipMemcpy(flags, h->d_wave, sizeof(flags), hipMemcpyDeviceToHost));
  return flags[1] ? icu_fail(ICU_E_STATE, "UMHS wavefront timed out") : ICU_OK;
}

int icu_umhs_estimate(icu_umhs* h, const int16_t* cur, const int16_t* ref, const int16_t* prev_mv,
                      const uint8_t* prev_intra, const uint8_t* include, int16_t* mv, int32_t* dist, long* avg) {
  if (!h || !cur || !ref || !mv) return icu_fail(ICU_E_ARG, "null argument");
  if ((prev_mv == nullptr) != (prev_intra == nullptr))
    return icu_fail(ICU_E_ARG, "previous vectors and intra flags go together");
  const size_t np = (size_t)h->W * h->H, nmb = (size_t)h->mbw * h->mbh;
  h->h8.resize(2 * np);
  for (size_t i = 0; i < np; i++) {
    if (cur[i] < 0 || cur[i] > 255 || ref[i] < 0 || ref[i] > 255)
      return icu_fail(ICU_E_RANGE, "input sample outside 0..255");
    h->h8[i] = (uint8_t)cur[i];
    h->h8[np + i] = (uint8_t)ref[i];
  }
  UM_TRY(hipMemcpyAsync(h->d_cur, h->h8.data(), np, hipMemcpyHostToDevice, h->own));
  UM_TRY(hipMemcpyAsync(h->d_ref, h->h8.data() + np, np, hipMemcpyHostToDevice, h->own));
  if (prev_mv) {
    UM_TRY(hipMemcpyAsync(h->d_prev, prev_mv, 4 * nmb, hipMemcpyHostToDevice, h->own));
    UM_TRY(hipMemcpyAsync(h->d_prev_intra, prev_intra, nmb, hipMemcpyHostToDevice, h->own));
  }
  int rc = icu_umhs_estimate_device(h, h->d_cur, h->d_ref, prev_mv ? h->d_prev : nullptr,
                                    prev_mv ? h->d_prev_intra : nullptr, h->d_mv, h->d_dist, h->own);
  if (rc) return rc;
  std::vector<int32_t> dd(nmb);
  UM_TRY(hipMemcpyAsync(mv, h->d_mv, 4 * nmb, hipMemcpyDeviceToHost, h->own));
  UM_TRY(hipMemcpyAsync(dd.data(), h->d_dist, 4 * nmb, hipMemcpyDeviceToHost, h->own));
  // everything on the handle's own stream: objects driven from different
  // threads overlap on the device (a null-stream copy would wait for them all)
  int flags[2] = {0, 0};
  UM_TRY(hipMemcpyAsync(flags, h->d_wave, sizeof(flags), hipMemcpyDeviceToHost, h->own));
  UM_TRY(hipStreamSynchronize(h->own));
  if (flags[1]) return icu_fail(ICU_E_STATE, "UMHS wavefront timed out");
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
