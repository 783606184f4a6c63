// Per-macroblock decision, sub-pel refinement and final vector choice.
//
// One wavefront (64 lanes) decides one MB.  It restates, in order, the rest
// of MotionEstimatorH264ImplFull::Estimate
// (src/CodecUtils/MotionEstimatorH264ImplFull.cpp:350-601):
//   * the integer decision, either from the search kernel's packed argmin
//     (zero predictor, filter off) or by replaying the reference's scan over
//     the cost table: with the running-best filter (:410-411) the scan is
//     path dependent, so lanes test 64 consecutive candidates at once against
//     the current best, a ballot finds the first one the reference would
//     accept, and the scan restarts after it;
//   * the half-pel stage (LoadHalfQuartPelWindow :1095-1182, loop :452-483)
//     from a 35x35 grid of G/b/h/j samples built in LDS;
//   * the quarter-pel stage (LoadQuartPelWindow :1198-1222, loop :485-517)
//     with the estimator's own pairwise averages;
//   * the predictor distortion (:529-570) and GetBestVector
//     (MotionEstimatorH264ImplFull.h:140-168) in correctly rounded doubles.
// With the median predictor (H264MotionVectorPredictorImpl1 over
// MacroBlockH264::GetMbMotionMedianPred, MacroBlockH264.cpp:757-816) an MB
// depends on its left, above-left, above and above-right neighbours, so MB
// rows run as a wavefront: each row is owned by one wave that polls the row
// above's vector words until they are final.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "icu_decide.h"

namespace icu {

namespace {

using namespace dec;

constexpr int kSpinLimit = 1 << 26;

__global__ __launch_bounds__(64) void decide_parallel_kernel(const DecideDesc* __restrict__ descs,
                                                             const DecideItem* __restrict__ items,
                                                             int nitems) {
  __shared__ WaveLds s;
  const int lane = threadIdx.x;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const DecideItem item = items[it];
    decide_mb(descs[item.stream], s, item.mb, 0, 0, lane);
    __syncthreads();
  }
}

// The median-predictor mode's predictor-free pass, one wave per MB, all MBs
// in parallel: the integer stage without a predictor and the sub-pel
// refinement (predfree_mb).
__global__ __launch_bounds__(64) void decide_predfree_kernel(const DecideDesc* __restrict__ desc,
                                                             PredFree* __restrict__ pf, int nmb) {
  __shared__ WaveLds s;
  const int lane = threadIdx.x;
  const DecideDesc& d = *desc;
  // the wavefront that follows starts from row ticket 0 and polls each
  // neighbour's vector word until it no longer holds kVecPending
  if (blockIdx.x == 0 && lane == 0) d.ticket[0] = 0;
  for (int mb = blockIdx.x; mb < nmb; mb += gridDim.x) {
    if (lane == 0) *(uint32_t*)(d.mv_out + 2 * (d.mb_base + mb)) = kVecPending;
    predfree_mb(d, s, mb, lane, pf + mb);
    __syncthreads();
  }
}

// One wave per MB row; rows are handed out by ticket in start order, so the
// row a wave waits on always belongs to a wave that is already running.
// The hand-off is the vector word itself: the predictor-free pass sets every
// MB's word to kVecPending, an MB polls its above and above-right (or
// above-left) words until both are final, and its own single vector store
// publishes it.  No progress counter, and no drain of the other stores
// before publishing: one memory round trip per MB on the critical path
// instead of three.
// pf: the MBs' predictor-free passes; an MB whose scan met no tie only adds
// its predictor's distortion and the final choice, one that did is decided
// in full.  stage_tab: such an MB's cost table (table_dim^2 dwords) is
// copied into the dynamic LDS with global_load_lds before the MB waits for
// its neighbours, so the filtered scan's rounds read LDS instead of HBM.
__global__ __launch_bounds__(64) void decide_wavefront_kernel(const DecideDesc* __restrict__ desc, int stage_tab,
                                                              const PredFree* __restrict__ pf) {
  __shared__ WaveLds s;
  __shared__ int row_s;
  extern __shared__ uint32_t s_tab[];
  const DecideDesc& d = *desc;
  const int lane = threadIdx.x;
  int* ticket = d.ticket;
  if (lane == 0) row_s = atomicAdd(&ticket[0], 1);
  __syncthreads();
  const int row = row_s;
  if (row >= d.mbh) return;
  const int n2 = d.table_dim * d.table_dim;
  int Ax = 0, Ay = 0;  // the left neighbour: this wave's previous MB
  for (int col = 0; col < d.mbw; col++) {
    // what does not depend on the neighbours is read before waiting for them
    const PredFree mine = pf[row * d.mbw + col];
    const uint32_t cw =
        *(const uint32_t*)(d.cur + (long)(16 * row + (lane >> 2)) * d.cur_stride + 16 * col + 4 * (lane & 3));
    const bool stage = stage_tab && mine.tie;
    if (stage) {
      const uint32_t* src = d.table + d.table_base + (long long)(row * d.mbw + col) * n2;
      for (int base = 0; base < n2; base += 64)
        if (base + lane < n2)
          __builtin_amdgcn_global_load_lds((const void*)(src + base + lane),
                                           (__attribute__((address_space(3))) void*)(s_tab + base), 4, 0, 0);
    }
    // neighbour vectors (MacroBlockH264.cpp:757-816)
    const int mb = row * d.mbw + col;
    // each (x, y) pair is one aligned 32-bit word; read it coherently
    const uint32_t* mvw = (const uint32_t*)(d.mv_out + 2 * d.mb_base);
    int Bx = 0, By = 0, Cx = 0, Cy = 0;
    if (row > 0) {
      // lane 0: above; lane 1: above-right, or above-left in the last column
      const bool needC = col + 1 < d.mbw || col > 0;
      const int k = lane == 0 ? mb - d.mbw : (col + 1 < d.mbw ? mb - d.mbw + 1 : mb - d.mbw - 1);
      uint32_t w = 0;
      // Bounded: a wait that runs out sets the error flag (ticket[1]), which
      // stays set until the host reads it; every other wait then gives up at
      // once instead of spinning its own full limit.
      if (lane == 0 || (lane == 1 && needC)) {
        int spins = 0;
        while ((w = __hip_atomic_load(&mvw[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == kVecPending) {
          if (++spins > kSpinLimit) {
            __hip_atomic_store(&ticket[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          if ((spins & 63) == 0 && poll_progress(&ticket[1])) break;
          __builtin_amdgcn_s_sleep(1);
        }
      }
      const uint32_t wB = __builtin_amdgcn_readlane(w, 0), wC = __builtin_amdgcn_readlane(w, 1);
      Bx = (int)(int16_t)(wB & 0xFFFF);
      By = (int)(int16_t)(wB >> 16);
      if (needC) {
        Cx = (int)(int16_t)(wC & 0xFFFF);
        Cy = (int)(int16_t)(wC >> 16);
      }
    } else {
      Bx = Ax; By = Ay; Cx = Ax; Cy = Ay;
    }
    const int px = median3(Ax, Bx, Cx), py = median3(Ay, By, Cy);
    if (stage) {
      __builtin_amdgcn_s_waitcnt(0);  // the table has landed
      wave_sync();
    }
    const Decision dec = decide_mb_after(d, s, mb, px, py, lane, mine, cw, stage ? s_tab : nullptr);
    Ax = dec.mvx;
    Ay = dec.mvy;
    __syncthreads();
  }
}

__global__ void fill_i32_kernel(int* p, int v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void sqrt_raw_kernel(const int* m, double* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = sqrt_int_rn(m[i]);
}

}  // namespace

void launch_decide_parallel(const DecideDesc* d_descs, const DecideItem* d_items, int nitems,
                            hipStream_t st) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(decide_parallel_kernel, dim3(nitems), dim3(64), 0, st, d_descs, d_items, nitems);
}

void launch_decide_predfree(const DecideDesc* d_desc, void* d_pf, int nmb, hipStream_t st) {
  if (nmb <= 0) return;
  hipLaunchKernelGGL(decide_predfree_kernel, dim3(nmb), dim3(64), 0, st, d_desc, (PredFree*)d_pf, nmb);
}

void launch_decide_wavefront(const DecideDesc* d_desc, int table_dim, int mbh, const void* d_pf,
                             hipStream_t st) {
  // stage a tie MB's table in LDS when it fits 48 KB (ranges up to +-55)
  const size_t bytes = (size_t)table_dim * table_dim * 4;
  const int stage = table_dim > 0 && bytes <= 48 * 1024;
  hipLaunchKernelGGL(decide_wavefront_kernel, dim3(mbh), dim3(64), stage ? bytes : 0, st, d_desc, stage,
                     (const PredFree*)d_pf);
}

int predfree_bytes() { return (int)sizeof(PredFree); }

void launch_fill_i32(int* p, int v, int n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
}

void launch_sqrt_check(const int* m, double* out, int n, hipStream_t st) {
  hipLaunchKernelGGL(sqrt_raw_kernel, dim3((n + 255) / 256), dim3(256), 0, st, m, out, n);
}

}  // namespace icu
