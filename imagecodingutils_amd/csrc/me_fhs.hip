// The FHS fast estimator (icu_fhs_*): MotionEstimatorH264ImplFHS
// (src/CodecUtils/MotionEstimatorH264ImplFHS.cpp:497-905) on the device.
//
// Per MB, as the reference: the median predictor with distortion, early
// exits on the predictor and the zero vector, an uneven multi-hexagon
// (scales 1..4), one extended hexagon and one diamond, every full-pel
// distortion the 80-point path sum of Td16x16OptimalPathLessThan
// (:1685-1790, extrapolated to 256 points at the first checkpoint past the
// minimum), the integer cost (d + 269 floor(sqrt(dist^2))) >> 8, then
// HalfPelEstimate / QuarterPelEstimate (:1109-1489) and the final test
// against the predictor.  MB rows run as a wavefront (one wave per row).
//
// The half-pel stage fills a b / h / j sample cache that the reference keeps
// in the estimator: h (rows -1..15, cols 0..15), then h's side columns and b
// unless the minimum fell below 1000, then b's side rows and j unless it fell
// again.  The quarter stage reads the cache whatever was filled, so its
// samples can come from the last MB (in raster order, or in an earlier call)
// that filled that part.  Each part's values depend only on its filler's
// sub-pel origin and the reference, so a quarter stage rebuilds a part it
// did not fill from the half-pel lattice at its last filler's origin: an
// earlier MB of the same row (known to the wave), else of an earlier row
// (found once the row above is complete), else the previous call's final
// cache, which a one-wave kernel stores after each call.  Entries no call
// ever wrote read as 0 (the reference's are uninitialised) and are counted.
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

// The first 80 points of the sampling path (FHS.cpp:1685-1790) as y*16 + x.
__constant__ uint8_t kPath80[80] = {
    165, 93,  18,  237, 225, 25,  97,  159, 216, 46,  71,  138, 244, 160, 48,  5,
    188, 116, 12,  234, 195, 90,  223, 52,  135, 126, 214, 59,  146, 169, 31,  39,
    83,  248, 193, 104, 16,  156, 254, 79,  203, 102, 183, 112, 190, 7,   242, 124,
    50,  148, 44,  57,  220, 3,   229, 69,  208, 152, 178, 251, 26,  114, 111, 201,
    37,  197, 65,  91,  141, 121, 61,  133, 186, 246, 173, 72,  163, 29,  129, 14};
__constant__ int8_t kHex16[16][2] = {{-4, 0}, {4, 0},  {-4, 1}, {4, 1},  {-4, -1}, {4, -1}, {-4, 2}, {4, 2},
                                     {-4, -2}, {4, -2}, {-2, 3}, {2, 3}, {-2, -3}, {2, -3}, {0, 4},  {0, -4}};
__constant__ int8_t kExtHex[6][2] = {{-2, 0}, {-1, -2}, {1, -2}, {2, 0}, {1, 2}, {-1, 2}};
__constant__ int8_t kCross[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};

constexpr int kSpin = 1 << 26;
constexpr int kThr = 1000;  // MEH264IFHS_THRESHOLD_MIN

// Cache parts: h main, h side columns, b main, b side rows, j.
enum : int { kHm = 1, kHs = 2, kBm = 4, kBs = 8, kJ = 16, kAll = 31 };
constexpr int kCache = 3 * 18 * 18;

struct FhsDesc {
  DecideDesc d;
  int mrng, mode;
  int16_t* mv_out;
  int32_t* dist_out;
  int* progress;       // mbh counters
  int* flags;          // [0] row ticket, [1] timed out, [2] unwritten cache reads
  int* fill_mask;      // per MB: cache parts filled
  int* fill_org;       // per MB: sub-pel origin (x | y << 16)
  int16_t* carry;      // the cache after the previous call, [3][18][18]
  uint8_t* carry_w;    // entries written by some call
};

struct FhsLds {
  WaveLds w;
  int16_t fc[3][18][18];  // the reference's cache as the quarter stage sees it
  uint8_t fw[3][18][18];  // entry written (else it reads 0)
};

__device__ __forceinline__ int fhs_cost(int d, int dx, int dy) {
  return (d + 269 * (int)sqrt_int_rn(dx * dx + dy * dy)) >> 8;
}

// The 10 sums of 8 consecutive path points, from each lane's squared error
// at points lane (e1) and 64 + lane (e2, lanes 0..15).
struct Path80 {
  int g[10];
  // Td16x16OptimalPathLessThan's return value against min
  __device__ __forceinline__ int lt(int min) const {
    int dp = g[0] + g[1] + g[2] + g[3];
#pragma unroll
    for (int k = 4; k < 10; k++) {
      const int pd = (dp * 256) / (8 * k);
      if (pd > min) return pd;
      dp += g[k];
    }
    return (dp * 256) / 80;
  }
};

// (DPP row shifts by 1, 2 and 4 leave lane 8 g + 7 the sum of lanes
// 8 g .. 8 g + 7, every group lying within one 16-lane row)
__device__ __forceinline__ int scan8(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);  // row_shr:4
  return v;
}

template <int K>
__device__ __forceinline__ void path_sums(int (&e1)[K], int (&e2)[K], Path80 (&out)[K]) {
#pragma unroll
  for (int k = 0; k < K; k++) {
    e1[k] = scan8(e1[k]);
    e2[k] = scan8(e2[k]);
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
#pragma unroll
    for (int g = 0; g < 8; g++) out[k].g[g] = __builtin_amdgcn_readlane(e1[k], 8 * g + 7);
    out[k].g[8] = __builtin_amdgcn_readlane(e2[k], 7);
    out[k].g[9] = __builtin_amdgcn_readlane(e2[k], 15);
  }
}

__device__ __forceinline__ int sq(int a, int b) { return (a - b) * (a - b); }

struct Ctx {
  const FhsDesc& f;
  const DecideDesc& d;
  int n, m, lane;
  int cv1, cv2;              // input pels at this lane's path points
  const uint8_t* rp1;        // reference pel of path point lane at the zero vector
  const uint8_t* rp2;        // and of point 64 + lane
  int p1x, p1y, p2x, p2y;    // the two path points in the block
  const uint8_t* win;        // the MB's staged search window
  fw::Win w;
  int xl, xr, yu, yd, pX0, pY0;
  int minD, minC, rmx, rmy;
};

// Up to 8 candidates at offsets (ox, oy) from (mx, my) in the reference's
// loop body: a window test, the path distortion against the current
// minimum, no worse distortion and strictly lower cost.
__device__ __forceinline__ void try8(Ctx& c, int mx, int my, const int (&ox)[8], const int (&oy)[8], int cnt) {
  bool ok[8];
  int e1[8], e2[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int x = mx + ox[k], y = my + oy[k];
    ok[k] = k < cnt && y >= c.yu && y <= c.yd && x >= c.xl && x <= c.xr;
    // window candidates are always staged
    const int bx = c.n + x, by = c.m + y;
    e1[k] = ok[k] ? sq(c.cv1, fw::byte_at(c.win, c.w, bx + c.p1x, by + c.p1y)) : 0;
    e2[k] = (ok[k] && c.lane < 16) ? sq(c.cv2, fw::byte_at(c.win, c.w, bx + c.p2x, by + c.p2y)) : 0;
  }
  Path80 p[8];
  path_sums(e1, e2, p);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (!ok[k]) continue;
    const int v = p[k].lt(c.minD);
    if (v <= c.minD) {
      const int cost = fhs_cost(v, mx + ox[k] - c.pX0, my + oy[k] - c.pY0);
      if (cost < c.minC) {
        c.minD = v;
        c.minC = cost;
        c.rmx = ox[k];
        c.rmy = oy[k];
      }
    }
  }
}

// A stage's candidates one per lane (lane k the stage's k-th, offset (ox,
// oy) from (mx, my)): each lane sums its candidate's squared errors along
// the 80 path points in the ten groups of 8 the early exit looks at (the
// block's pels at the points from the lanes that hold them, the reference's
// from the staged window), then the reference's loop body is replayed in
// order by ballot: the first candidate after the last accepted one whose
// path distortion against the current minimum is no worse and whose cost
// is strictly lower replaces the best, until none is left.  One pass for up
// to 64 candidates, where try8 takes one per 8.  Lanes without a candidate
// (live false, or outside the window) read the zero vector's block and
// take no part.
__device__ __forceinline__ void try_lanes(Ctx& c, int mx, int my, int ox, int oy, bool live) {
  const int x = mx + ox, y = my + oy;
  const bool ok = live && y >= c.yu && y <= c.yd && x >= c.xl && x <= c.xr;
  const int bx = ok ? c.n + x : c.n, by = ok ? c.m + y : c.m;  // window candidates are always staged
  const uint8_t* base = c.win + (by - c.w.y0) * c.w.cols + (bx - c.w.x0);
  Path80 p;
#pragma unroll
  for (int g = 0; g < 10; g++) {
    int a = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int t = 8 * g + i;
      const int pt = kPath80[t];
      const int cv = __builtin_amdgcn_readlane(t < 64 ? c.cv1 : c.cv2, t & 63);
      a += sq(cv, base[(pt >> 4) * c.w.cols + (pt & 15)]);
    }
    p.g[g] = a;
  }
  const int q = 269 * (int)sqrt_int_rn((x - c.pX0) * (x - c.pX0) + (y - c.pY0) * (y - c.pY0));
  int start = 0;
  for (;;) {
    const int v = p.lt(c.minD);
    const int cost = (v + q) >> 8;
    const uint64_t take = __ballot(ok && c.lane >= start && v <= c.minD && cost < c.minC);
    if (take == 0) break;
    const int k = __builtin_ctzll(take);
    c.minD = __builtin_amdgcn_readlane(v, k);
    c.minC = __builtin_amdgcn_readlane(cost, k);
    c.rmx = __builtin_amdgcn_readlane(ox, k);
    c.rmy = __builtin_amdgcn_readlane(oy, k);
    start = k + 1;
  }
}

__device__ __forceinline__ void try_table(Ctx& c, const int8_t (*tab)[2], int cnt, int w, int mx, int my) {
  for (int g = 0; g < cnt; g += 8) {
    int ox[8], oy[8];
    const int k8 = min(8, cnt - g);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int t = k < k8 ? g + k : g;
      ox[k] = w * tab[t][0];
      oy[k] = w * tab[t][1];
    }
    try8(c, mx, my, ox, oy, k8);
  }
}

// Copy cache parts `parts` from the lattice built at their filler's origin
// (s.w.Hg: G at (2 + 2r, 2 + 2c) for block pel (r, c)).
__device__ void lattice_to_cache(FhsLds& s, int parts, int lane) {
  for (int i = lane; i < kCache; i += 64) {
    const int t = i / 324, rr = (i / 18) % 18, cc = i % 18;
    int part;
    int v = 0;
    if (t == 0) {  // h(c, r) at row index r + 1 (r -1..15), col c + 1 (c -1..16)
      part = (cc == 0 || cc == 17) ? kHs : kHm;
      if (rr > 16) continue;
      v = s.w.Hg[(3 + 2 * (rr - 1)) * kHalf + 2 + 2 * (cc - 1)];
    } else if (t == 1) {  // b(c, r): rows r -1..16, cols c -1..15
      part = (rr == 0 || rr == 17) ? kBs : kBm;
      if (cc > 16) continue;
      v = s.w.Hg[(2 + 2 * (rr - 1)) * kHalf + 3 + 2 * (cc - 1)];
    } else {  // j(c, r): r, c -1..15
      part = kJ;
      if (rr > 16 || cc > 16) continue;
      v = s.w.Hg[(3 + 2 * (rr - 1)) * kHalf + 3 + 2 * (cc - 1)];
    }
    if (!(parts & part)) continue;
    s.fc[t][rr][cc] = (int16_t)v;
    s.fw[t][rr][cc] = 1;
  }
}

__device__ __forceinline__ int part_of(int t, int rr, int cc) {
  if (t == 0) return (cc == 0 || cc == 17) ? kHs : kHm;
  if (t == 1) return (rr == 0 || rr == 17) ? kBs : kBm;
  return kJ;
}

__device__ void carry_to_cache(const FhsDesc& f, FhsLds& s, int parts, int lane) {
  for (int i = lane; i < kCache; i += 64) {
    const int t = i / 324, rr = (i / 18) % 18, cc = i % 18;
    if (!(parts & part_of(t, rr, cc))) continue;
    s.fc[t][rr][cc] = f.carry[i];
    s.fw[t][rr][cc] = f.carry_w[i];
  }
}

__device__ __forceinline__ void build_lattice(const DecideDesc& d, FhsLds& s, int org, int lane,
                                              const uint8_t* win = nullptr, const fw::Win* w = nullptr) {
  const int x0 = (int)(int16_t)(org & 0xFFFF), y0 = org >> 16;
  if (win) fw::patch(d, win, *w, w->has(x0, y0, 16, 16, 3), x0, y0, s.w.P, lane);
  else load_patch(d, x0, y0, s.w.P, lane);
  wave_sync();
  build_half_grid(s.w, lane);
}

__device__ __forceinline__ int pack_org(int x, int y) { return (int)(((uint32_t)x & 0xFFFFu) | ((uint32_t)y << 16)); }

__global__ __launch_bounds__(64) void fhs_kernel(FhsDesc f) {
  __shared__ FhsLds s;
  __shared__ __align__(16) uint8_t s_win[fw::kBytes];
  __shared__ int row_s;
  const DecideDesc& d = f.d;
  const int lane = threadIdx.x;
  if (lane == 0) row_s = atomicAdd(&f.flags[0], 1);
  __syncthreads();
  const int row = row_s;
  if (row >= d.mbh) return;
  const uint32_t* mvw = (const uint32_t*)f.mv_out;
  int row_org[5], row_has = 0;  // last filler of each cache part earlier in this row
#pragma unroll
  for (int i = 0; i < 5; i++) row_org[i] = 0;
  bool above_done = row == 0;
  for (int col = 0; col < d.mbw; col++) {
    const int n = col * 16, m = row * 16;
    // the MB's search window into LDS and its block into registers, while
    // the MB waits for its neighbours
    const fw::Win w = fw::shape(n, m, d.W, d.H, f.mrng - 1);
    fw::issue(d, w, s_win, lane);
    const uint32_t cw = *(const uint32_t*)(d.cur + (long)(m + (lane >> 2)) * d.cur_stride + n + 4 * (lane & 3));
    const int p1 = kPath80[lane], p2 = kPath80[64 + (lane & 15)];
    const int cv1 = d.cur[(long)(m + (p1 >> 4)) * d.cur_stride + n + (p1 & 15)];
    const int cv2 = d.cur[(long)(m + (p2 >> 4)) * d.cur_stride + n + (p2 & 15)];
    if (row > 0) {
      const int need = min(col + 2, d.mbw);
      if (lane == 0) {
        int spins = 0;
        while (poll_progress(&f.progress[row - 1]) < need) {
          if (++spins > kSpin) {
            atomicExch(&f.flags[1], 1);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __syncthreads();
    }
    fw::wait();
    const int mb = row * d.mbw + col;
    // ---- predictor with distortion (MacroBlockH264.cpp:818-915)
    auto ld = [&](int k, int* x, int* y, int* dd) {
      const uint32_t w = __hip_atomic_load(&mvw[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *x = (int)(int16_t)(w & 0xFFFF);
      *y = (int)(int16_t)(w >> 16);
      *dd = __hip_atomic_load(&f.dist_out[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    int Ax = 0, Ay = 0, Ad = 0, Bx = 0, By = 0, Bd = 0, Cx = 0, Cy = 0, Cd = 0;
    const bool hasB = row > 0, hasC = row > 0 && col + 1 < d.mbw;
    if (col > 0) ld(mb - 1, &Ax, &Ay, &Ad);
    if (hasB) ld(mb - d.mbw, &Bx, &By, &Bd);
    if (hasC) ld(mb - d.mbw + 1, &Cx, &Cy, &Cd);
    else if (col > 0 && row > 0) ld(mb - d.mbw - 1, &Cx, &Cy, &Cd);
    if (!hasB && !hasC) {
      Bx = Cx = Ax;
      By = Cy = Ay;
      Bd = Cd = Ad;
    }
    const int predX = median3(Ax, Bx, Cx), predY = median3(Ay, By, Cy);
    int xd = Ad, yd = Ad;
    if (predX == Bx && Bd < xd) xd = Bd;
    if (predX == Cx && Cd < xd) xd = Cd;
    if (predY == By && Bd < yd) yd = Bd;
    if (predY == Cx && Cd < yd) yd = Cd;  // the reference compares the y median with C's x
    const int predD = yd < xd ? yd : xd;

    // ---- the block and this lane's path points
    s.w.cur4[lane] = cw;
    Ctx c{f, d, n, m, lane};
    c.cv1 = cv1;
    c.cv2 = cv2;
    c.rp1 = d.ref + (long)(m + (p1 >> 4)) * d.ref_stride + n + (p1 & 15);
    c.rp2 = d.ref + (long)(m + (p2 >> 4)) * d.ref_stride + n + (p2 & 15);
    c.p1x = p1 & 15;
    c.p1y = p1 >> 4;
    c.p2x = p2 & 15;
    c.p2y = p2 >> 4;
    c.win = s_win;
    c.w = w;
    wave_sync();

    int pXQ = predX % 4, pYQ = predY % 4;
    int predX0 = predX / 4, predY0 = predY / 4;
    const int oX0 = predX0, oY0 = predY0;
    if (predX0 + n > d.W) predX0 = 16;
    if (predX0 + n < -16) predX0 = -16;
    if (predY0 + m > d.H) predY0 = 16;
    if (predY0 + m < -16) predY0 = -16;
    if (oX0 != predX0 || oY0 != predY0) pXQ = pYQ = 0;
    const int rX = predX0 * 4 + pXQ, rY = predY0 * 4 + pYQ;
    motion_window(n, m, d.W, d.H, f.mrng - 1, &c.xl, &c.xr, &c.yu, &c.yd);
    c.pX0 = predX0;
    c.pY0 = predY0;
    int predVecD;
    if (pXQ || pYQ) {
      // QuarterRead at the predictor, filtered in registers (SSD)
      predVecD = quarter_dist(d, s.w, n + predX0, m + predY0, pXQ, pYQ, lane);
    } else {
      const int px = n + predX0, py = m + predY0;
      const uint32_t rw = fw::row4(d, s_win, w, w.has(px, py, 16, 16, 0), px, py, lane);
      int a = 0;
#pragma unroll
      for (int i = 0; i < 4; i++) a += sq((int)((cw >> (8 * i)) & 0xFFu), (int)((rw >> (8 * i)) & 0xFFu));
      predVecD = wave_sum(a);
    }
    c.minD = predVecD;
    c.minC = predVecD / 256;
    c.rmx = c.rmy = 0;
    int mx = predX0, my = predY0, mvx, mvy;
    int filled = 0, org = 0;
    bool done = false;
    if (predVecD < kThr) {
      mvx = rX;
      mvy = rY;
      done = true;
    }
    if (!done && (rX || rY)) {
      int e1[1] = {sq(c.cv1, fw::byte_at(s_win, w, n + c.p1x, m + c.p1y))};
      int e2[1] = {lane < 16 ? sq(c.cv2, fw::byte_at(s_win, w, n + c.p2x, m + c.p2y)) : 0};
      Path80 p[1];
      path_sums(e1, e2, p);
      const int zd = p[0].lt(c.minD);
      if (zd < c.minD) {
        const int zc = fhs_cost(zd, -predX0, -predY0);
        if (zc < c.minC) {
          c.minD = zd;
          c.minC = zc;
          mx = my = 0;
        }
        if (zd < kThr) {
          mvx = mvy = 0;
          done = true;
        }
      }
    }
    if (!done) {
      int stage;
      if (c.minD < kThr) stage = 2;
      else if (c.minD < 4000 || (c.minD == predVecD && c.minD < predD * 12 / 10 && c.minD > predD * 8 / 10))
        stage = 1;
      else stage = 0;
      if (stage == 0) {
        c.rmx = c.rmy = 0;
        {
          // the four scales' 16 points each, scale by scale: 64 candidates
          const int k = lane & 15, w = (lane >> 4) + 1;
          try_lanes(c, mx, my, w * kHex16[k][0], w * kHex16[k][1], true);
        }
        mx += c.rmx;
        my += c.rmy;
      }
      if (stage <= 1) {
        c.rmx = c.rmy = 0;
        try_table(c, kExtHex, 6, 1, mx, my);
        mx += c.rmx;
        my += c.rmy;
      }
      c.rmx = c.rmy = 0;
      try_table(c, kCross, 4, 1, mx, my);
      mx += c.rmx;
      my += c.rmy;
      int minD = c.minD;
      mvx = mx * 4;
      mvy = my * 4;
      if (f.mode <= 1) {
        // ---- HalfPelEstimate: the lattice at the full-pel winner, the 8
        // candidates summed at once, then taken in the reference's passes
        org = pack_org(n + mx, m + my);
        build_lattice(d, s, org, lane, s_win, &w);
        const uint8_t* cur8 = (const uint8_t*)s.w.cur4;
        int acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int p = lane; p < 256; p += 64) {
          const int r = p >> 4, cc = p & 15;
          const int cvv = cur8[p];
          const int16_t* h = s.w.Hg + (2 + 2 * r) * kHalf + 2 + 2 * cc;
#pragma unroll
          for (int k = 0; k < 8; k++) acc[k] += sq(cvv, h[kSubY[k] * kHalf + kSubX[k]]);
        }
        wave_sum8(acc);
        int lmin = minD, hx = 0, hy = 0;
        filled = kHm;
        // h pass: (0,-1), (0,1)
        if (acc[1] < lmin) { lmin = acc[1]; hx = 0; hy = -1; }
        if (acc[6] < lmin) { lmin = acc[6]; hx = 0; hy = 1; }
        if (lmin >= kThr) {
          filled |= kHs | kBm;
          // b pass: (-1,0), (1,0)
          if (acc[3] < lmin) { lmin = acc[3]; hx = -1; hy = 0; }
          if (acc[4] < lmin) { lmin = acc[4]; hx = 1; hy = 0; }
          if (lmin >= kThr) {
            filled |= kBs | kJ;
            // j pass: (-1,-1), (1,-1), (-1,1), (1,1)
            if (acc[0] < lmin) { lmin = acc[0]; hx = -1; hy = -1; }
            if (acc[2] < lmin) { lmin = acc[2]; hx = 1; hy = -1; }
            if (acc[5] < lmin) { lmin = acc[5]; hx = -1; hy = 1; }
            if (acc[7] < lmin) { lmin = acc[7]; hx = 1; hy = 1; }
          }
        }
        if (f.mode == 1) {
          if (lmin < minD) {
            minD = lmin;
            mvx += 2 * hx;
            mvy += 2 * hy;
          }
        } else {
          // ---- QuarterPelEstimate
          int qmin = minD, qx = 0, qy = 0;
          bool qdone = false;
          if (lmin < minD) {
            qmin = lmin;
            if (lmin < kThr) qdone = true;
          }
          if (!qdone) {
            // the cache as the reference's quarter stage sees it
            lattice_to_cache(s, filled, lane);
            int missing = kAll & ~filled;
            for (int part = 0; part < 5 && missing; part++) {
              const int bit = 1 << part;
              if (!(missing & bit) || !(row_has & bit)) continue;
              const int o = row_org[part];
              // every part this row's filler at o also covers, at once
              int parts = 0;
#pragma unroll
              for (int q = 0; q < 5; q++)
                if ((missing & (1 << q)) && (row_has & (1 << q)) && row_org[q] == o) parts |= 1 << q;
              wave_sync();
              build_lattice(d, s, o, lane, s_win, &w);
              lattice_to_cache(s, parts, lane);
              missing &= ~parts;
            }
            if (missing) {
              // earlier rows: wait for the row above to complete, then the
              // last MB before this row that filled each part
              if (!above_done && row > 0) {
                if (lane == 0) {
                  int spins = 0;
                  while (poll_progress(&f.progress[row - 1]) <
                         d.mbw) {
                    if (++spins > kSpin) {
                      atomicExch(&f.flags[1], 1);
                      break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                  }
                }
                __syncthreads();
                above_done = true;
              }
              for (int k = row * d.mbw - 1; k >= 0 && missing; k--) {
                const int fm = __hip_atomic_load(&f.fill_mask[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const int parts = fm & missing;
                if (!parts) continue;
                const int o = __hip_atomic_load(&f.fill_org[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                wave_sync();
                build_lattice(d, s, o, lane, s_win, &w);
                lattice_to_cache(s, parts, lane);
                missing &= ~parts;
              }
              if (missing) carry_to_cache(f, s, missing, lane);
            }
            wave_sync();
            // the 8 quarter candidates around (2hx, 2hy): H.264 quarter
            // samples from the cache and the full pels (FHS.cpp:1240-1349)
            int qa[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            int unw = 0;
            const int x0 = n + mx, y0 = m + my;
            // full pels of the quarter samples: one pel left / above to one
            // right / below of the block
            const bool gst = w.has(x0, y0, 16, 16, 1);
            for (int p = lane; p < 256; p += 64) {
              const int r = p >> 4, cc = p & 15;
              const int cvv = cur8[p];
#pragma unroll
              for (int k = 0; k < 8; k++) {
                const int ox = 2 * hx + kSubX[k], oy = 2 * hy + kSubY[k];
                const int sx = ox < 0 ? -1 : 0, sy = oy < 0 ? -1 : 0;
                const int fx = ox - 4 * sx, fy = oy - 4 * sy;
                const int ci = sy + 1 + r, cj = sx + 1 + cc;
                const uint8_t* gp = gst ? s_win + (y0 + sy + r - w.y0) * w.cols + (x0 + sx + cc - w.x0)
                                        : d.ref + (long)(y0 + sy + r) * d.ref_stride + x0 + sx + cc;
                auto C = [&](int t, int a, int b) {
                  unw += !s.fw[t][a][b];
                  return (int)s.fc[t][a][b];
                };
                int v;
                switch (fy * 4 + fx) {
                  case 1: v = (gp[0] + C(1, ci, cj) + 1) >> 1; break;
                  case 3: v = (gp[1] + C(1, ci, cj) + 1) >> 1; break;
                  case 4: v = (gp[0] + C(0, ci, cj) + 1) >> 1; break;
                  case 12: v = (gp[gst ? w.cols : d.ref_stride] + C(0, ci, cj) + 1) >> 1; break;
                  case 5: v = (C(1, ci, cj) + C(0, ci, cj) + 1) >> 1; break;
                  case 7: v = (C(1, ci, cj) + C(0, ci, cj + 1) + 1) >> 1; break;
                  case 6: v = (C(1, ci, cj) + C(2, ci, cj) + 1) >> 1; break;
                  case 14: v = (C(1, ci + 1, cj) + C(2, ci, cj) + 1) >> 1; break;
                  case 9: v = (C(0, ci, cj) + C(2, ci, cj) + 1) >> 1; break;
                  case 11: v = (C(0, ci, cj + 1) + C(2, ci, cj) + 1) >> 1; break;
                  case 13: v = (C(0, ci, cj) + C(1, ci + 1, cj) + 1) >> 1; break;
                  default: v = (C(0, ci, cj + 1) + C(1, ci + 1, cj) + 1) >> 1; break;  // 15
                }
                qa[k] += sq(cvv, v);
              }
            }
            wave_sum8(qa);
            // reads of entries no call wrote count only for candidates the
            // loop reaches; the count is a diagnostic, summed over all eight
            unw = wave_sum(unw);
            if (lane == 0 && unw) atomicAdd(&f.flags[2], 1);
#pragma unroll
            for (int k = 0; k < 8; k++) {
              if (qa[k] < qmin) {
                qmin = qa[k];
                qx = 2 * hx + kSubX[k];
                qy = 2 * hy + kSubY[k];
                if (qmin < kThr) break;
              }
            }
          }
          if (qmin < minD) {
            minD = qmin;
            mvx += qx;
            mvy += qy;
          }
        }
        wave_sync();
      }
      if (fhs_cost(predVecD, rX - predX0, rY - predY0) <= fhs_cost(minD, mvx - predX0, mvy - predY0)) {
        minD = predVecD;
        mvx = rX;
        mvy = rY;
      }
      c.minD = minD;
    }
    // record this MB's cache fills for later MBs
    if (f.mode == 0 && filled) {
#pragma unroll
      for (int q = 0; q < 5; q++)
        if (filled & (1 << q)) row_org[q] = org;
      row_has |= filled;
    }
    if (lane == 0) {
      __hip_atomic_store(&f.fill_mask[mb], f.mode == 0 ? filled : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&f.fill_org[mb], org, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t w = (uint32_t)(uint16_t)(int16_t)mvx | ((uint32_t)(uint16_t)(int16_t)mvy << 16);
      __hip_atomic_store((uint32_t*)(f.mv_out + 2 * mb), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&f.dist_out[mb], c.minD, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      publish_progress(&f.progress[row], col + 1);
    }
    __syncthreads();
  }
}

// After a call: the cache as the reference leaves it, each part from its
// last filler in the picture (parts no MB filled keep the earlier values).
__global__ __launch_bounds__(64) void fhs_carry_kernel(FhsDesc f) {
  __shared__ FhsLds s;
  const DecideDesc& d = f.d;
  const int lane = threadIdx.x;
  for (int i = lane; i < kCache; i += 64) {
    s.fc[i / 324][(i / 18) % 18][i % 18] = f.carry[i];
    s.fw[i / 324][(i / 18) % 18][i % 18] = f.carry_w[i];
  }
  int missing = kAll;
  for (int k = d.mbw * d.mbh - 1; k >= 0 && missing; k--) {
    const int parts = f.fill_mask[k] & missing;
    if (!parts) continue;
    wave_sync();
    build_lattice(d, s, f.fill_org[k], lane);
    lattice_to_cache(s, parts, lane);
    missing &= ~parts;
  }
  wave_sync();
  for (int i = lane; i < kCache; i += 64) {
    f.carry[i] = s.fc[i / 324][(i / 18) % 18][i % 18];
    f.carry_w[i] = s.fw[i / 324][(i / 18) % 18][i % 18];
  }
}

}  // namespace

struct icu_fhs {
  int W = 0, H = 0, mbw = 0, mbh = 0, qrange = 0, mode = 0;
  uint8_t* pad_base = nullptr;
  uint8_t* pad = nullptr;
  int stride = 0, pad_wgs = 0;
  int* d_wave = nullptr;   // mbh progress counters, then ticket, timeout, unwritten reads
  int* d_fill = nullptr;   // per MB: fill mask, then fill origin
  int16_t* d_carry = nullptr;
  uint8_t* d_carry_w = nullptr;
  uint8_t *d_cur = nullptr, *d_ref = nullptr;
  int16_t* d_mv = nullptr;
  int32_t* d_dist = nullptr;
  std::vector<uint8_t> h8;
  hipStream_t own = nullptr;
  ~icu_fhs() {
    hipFree(pad_base);
    hipFree(d_wave);
    hipFree(d_fill);
    hipFree(d_carry);
    hipFree(d_carry_w);
    hipFree(d_cur);
    hipFree(d_ref);
    hipFree(d_mv);
    hipFree(d_dist);
    if (own) hipStreamDestroy(own);
  }
};

namespace {

#define FH_TRY(expr)                                                                                    \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess) return icu_fail(ICU_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

extern "C" {

int icu_fhs_create(int width, int height, int motion_range, int mode, icu_fhs** out) {
  if (!out) return icu_fail(ICU_E_ARG, "null argument");
  if (width < 16 || height < 16 || width % 16 || height % 16)
    return icu_fail(ICU_E_ARG, "width and height must be positive multiples of 16");
  if (width > 32752 || height > 32752) return icu_fail(ICU_E_ARG, "width and height must be below 32752");
  if (motion_range < 8 || motion_range > 4 * 64)
    return icu_fail(ICU_E_ARG, "motion range (quarter pels) must be 8..256");
  if (mode < 0 || mode > 2) return icu_fail(ICU_E_ARG, "mode must be 0 (quarter), 1 (half) or 2 (full pel)");
  icu_fhs* h = new icu_fhs();
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
      hipMalloc(&h->d_wave, sizeof(int) * (h->mbh + 3)) != hipSuccess ||
      hipMalloc(&h->d_fill, sizeof(int) * 2 * n) != hipSuccess ||
      hipMalloc(&h->d_carry, sizeof(int16_t) * kCache) != hipSuccess ||
      hipMalloc(&h->d_carry_w, kCache) != hipSuccess || hipMalloc(&h->d_cur, (size_t)width * height) != hipSuccess ||
      hipMalloc(&h->d_ref, (size_t)width * height) != hipSuccess || hipMalloc(&h->d_mv, 4 * n) != hipSuccess ||
      hipMalloc(&h->d_dist, 4 * n) != hipSuccess || hipStreamCreate(&h->own) != hipSuccess ||
      hipMemset(h->d_carry, 0, sizeof(int16_t) * kCache) != hipSuccess ||
      hipMemset(h->d_carry_w, 0, kCache) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    delete h;
    return icu_fail(ICU_E_DEVICE, "device allocation failed");
  }
  h->pad = h->pad_base + (size_t)kPad * h->stride + kPad;
  *out = h;
  return ICU_OK;
}

int icu_fhs_destroy(icu_fhs* h) {
  delete h;
  return ICU_OK;
}

int icu_fhs_estimate_device(icu_fhs* h, const uint8_t* d_cur, const uint8_t* d_ref, int16_t* d_mv, int32_t* d_dist,
                            void* stream) {
  if (!h || !d_cur || !d_ref || !d_mv || !d_dist) return icu_fail(ICU_E_ARG, "null argument");
  if (((uintptr_t)d_cur & 15) || ((uintptr_t)d_mv & 3) || ((uintptr_t)d_dist & 3))
    return icu_fail(ICU_E_ARG, "d_cur must be 16-byte aligned, vectors and distortions 4-byte aligned");
  const hipStream_t st = stream ? (hipStream_t)stream : h->own;
  PadSet ps{};
  ps.n = 1;
  ps.d[0] = PadDesc{d_ref, h->pad, h->W, h->H, h->W, h->stride, 0};
  launch_pad_set(ps, h->pad_wgs, st);
  launch_fill_i32(h->d_wave, 0, h->mbh + 3, st);
  FhsDesc f{};
  DecideDesc& d = f.d;
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
  d.subpel = 2;
  f.mrng = h->qrange / 4;
  f.mode = h->mode;
  f.mv_out = d_mv;
  f.dist_out = d_dist;
  f.progress = h->d_wave;
  f.flags = h->d_wave + h->mbh;
  f.fill_mask = h->d_fill;
  f.fill_org = h->d_fill + h->mbw * h->mbh;
  f.carry = h->d_carry;
  f.carry_w = h->d_carry_w;
  hipLaunchKernelGGL(fhs_kernel, dim3(h->mbh), dim3(64), 0, st, f);
  if (h->mode == 0) hipLaunchKernelGGL(fhs_carry_kernel, dim3(1), dim3(64), 0, st, f);
  FH_TRY(hipGetLastError());
  return ICU_OK;
}

int icu_fhs_check(icu_fhs* h, int* unwritten_reads) {
  if (!h) return icu_fail(ICU_E_ARG, "null argument");
  int flags[3] = {0, 0, 0};
  FH_TRY(hipDeviceSynchronize());
  FH_TRY(hipMemcpy(flags, h->d_wave + h->mbh, sizeof(flags), hipMemcpyDeviceToHost));
  if (unwritten_reads) *unwritten_reads = flags[2];
  return flags[1] ? icu_fail(ICU_E_STATE, "FHS wavefront timed out") : ICU_OK;
}

int icu_fhs_estimate(icu_fhs* h, const int16_t* cur, const int16_t* ref, const uint8_t* include, int16_t* mv,
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
  FH_TRY(hipMemcpyAsync(h->d_cur, h->h8.data(), np, hipMemcpyHostToDevice, h->own));
  FH_TRY(hipMemcpyAsync(h->d_ref, h->h8.data() + np, np, hipMemcpyHostToDevice, h->own));
  int rc = icu_fhs_estimate_device(h, h->d_cur, h->d_ref, h->d_mv, h->d_dist, h->own);
  if (rc) return rc;
  std::vector<int32_t> dd(nmb);
  FH_TRY(hipMemcpyAsync(mv, h->d_mv, 4 * nmb, hipMemcpyDeviceToHost, h->own));
  FH_TRY(hipMemcpyAsync(dd.data(), h->d_dist, 4 * nmb, hipMemcpyDeviceToHost, h->own));
  // everything on the handle's own stream: objects driven from different
  // threads overlap on the device (a null-stream copy would wait for them all)
  int flags[2] = {0, 0};
  FH_TRY(hipMemcpyAsync(flags, h->d_wave + h->mbh, sizeof(flags), hipMemcpyDeviceToHost, h->own));
  FH_TRY(hipStreamSynchronize(h->own));
  if (flags[1]) return icu_fail(ICU_E_STATE, "FHS wavefront timed out");
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
