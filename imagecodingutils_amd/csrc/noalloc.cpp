// One-shot estimation of a frame pair that allocates nothing: the variant of
// icu_me_estimate_device for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, hosts that forbid it).
//
// Everything the handle path keeps in its own allocations (the padded
// reference plane, the stream / decision / padding descriptors and the task
// list) lives in one file-scope device array, `scratch`, of kScratchBytes per
// GPU, held until the process ends.  Its device address is looked up once
// per device with hipGetSymbolAddress (host code never passes the array's
// name itself to a kernel).  A first kernel writes the descriptors (passed
// to it by value) and the task list into the scratch; then the padding and
// search kernels of the handle path run on it.  Nothing in the scratch has
// to start at zero: every call rewrites what it reads.  Every call on a GPU
// shares that GPU's scratch, so calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/icu_hip.h"
#include "icu_device.h"
#include "icu_internal.h"

namespace icu {
namespace {

constexpr size_t kScratchBytes = size_t(20) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
uint8_t* g_scratch[kMaxDevices];  // device address of `scratch` per device

// The current device's scratch address, looked up on the first call on
// that device (with that device current).
int device_scratch(uint8_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return icu_fail(ICU_E_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return icu_fail(ICU_E_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return icu_fail(ICU_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = (uint8_t*)p;
  }
  *out = g_scratch[dev];
  return ICU_OK;
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// Descriptors into the scratch, and the task list in the order the handle
// path sorts it: one MB per task for the wave kernel (G = 0), else the full
// tasks (G MBs) row by row, then each row's shorter last task.
__global__ __launch_bounds__(256) void setup_kernel(const StreamDesc sd, const DecideDesc dd, const PadDesc pd,
                                                    StreamDesc* o_sd, DecideDesc* o_dd, PadDesc* o_pd,
                                                    SearchTask* tasks, int ntasks, int mbw, int G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    *o_sd = sd;
    *o_dd = dd;
    *o_pd = pd;
  }
  if (i >= ntasks) return;
  SearchTask t;
  t.stream = 0;
  if (G == 0) {
    t.mb = i;
    t.count = 1;
    tasks[i] = t;
    return;
  }
  const int per = mbw / G, nfull = per * sd.mbh;
  if (i < nfull) {
    const int row = i / per;
    t.mb = row * mbw + (i - row * per) * G;
    t.count = G;
  } else {
    t.mb = (i - nfull) * mbw + per * G;
    t.count = mbw - per * G;
  }
  tasks[i] = t;
}

}  // namespace
}  // namespace icu

using namespace icu;

extern "C" int icu_me_estimate_device_noalloc(const icu_me_config* cfg, const uint8_t* d_cur, const uint8_t* d_ref,
                                              int16_t* d_mv, int32_t* d_dist, int32_t* d_int, void* stream) {
  if (!cfg || !d_cur || !d_ref || !d_mv || !d_dist) return icu_fail(ICU_E_ARG, "null argument");
  std::string why;
  if (!valid_cfg(*cfg, &why)) return icu_fail(ICU_E_ARG, why);
  if (cfg->filter || cfg->predictor)
    return icu_fail(ICU_E_ARG, "the running-best filter and the median predictor keep a per-MB cost table; "
                               "the no-allocation estimate takes filter 0 and predictor 0");
  if ((uintptr_t)d_cur & 15) return icu_fail(ICU_E_ARG, "device current planes must be 16-byte aligned");
  const int W = cfg->width, H = cfg->height, mbw = W / 16, mbh = H / 16;
  const int L = cfg->window == ICU_WINDOW_FULL ? cfg->range : cfg->range - 1;
  const Tiling t = plan_tiling(mbw, L, cfg->metric, true, cfg->tile_rows, cfg->mbs_per_task);
  if (t.V == 0) return icu_fail(ICU_E_ARG, "tile_rows too tall for exact SAD argmin keys at this range");
  // one MB per wave where the window takes it (the handle path's default)
  const int L0 = cfg->metric == ICU_METRIC_SAD && (cfg->grid_mode == 0 || cfg->grid_mode >= 4) ? wave_class(L) : 0;
  const int per = mbw / t.G, ntasks = L0 ? mbw * mbh : per * mbh + (mbw % t.G ? mbh : 0);
  // scratch layout: descriptors, tasks, padded reference plane
  const int ref_stride = (W + 2 * kPad + 255) & ~255;
  const size_t o_dec = up256(sizeof(StreamDesc)), o_pad = o_dec + up256(sizeof(DecideDesc));
  const size_t o_task = o_pad + up256(sizeof(PadDesc));
  const size_t o_plane = o_task + up256(sizeof(SearchTask) * (size_t)ntasks);
  const size_t need = o_plane + (size_t)ref_stride * (H + 2 * kPad);
  if (need > kScratchBytes)
    return icu_fail(ICU_E_RANGE, "picture needs " + std::to_string(need) + " bytes of scratch; the no-allocation "
                                 "estimate has " + std::to_string(kScratchBytes) + " (up to about 3840x2160 "
                                 "fits)");
  uint8_t* base = nullptr;
  int rc = device_scratch(&base);
  if (rc) return rc;
  const hipStream_t st = (hipStream_t)stream;
  uint8_t* refpad = base + o_plane + (size_t)kPad * ref_stride + kPad;

  StreamDesc sd{};
  sd.cur = d_cur;
  sd.ref = refpad;
  sd.cur_stride = W;
  sd.ref_stride = ref_stride;
  sd.W = W;
  sd.H = H;
  sd.mbw = mbw;
  sd.mbh = mbh;
  sd.R = cfg->range;
  sd.L = L;
  sd.V = t.V;
  sd.B = t.B;
  sd.Q = t.Q;
  sd.G = t.G;
  sd.metric = cfg->metric;
  sd.mb_base = 0;
  sd.table_base_valid = 0;
  sd.table_base = 0;
  sd.table_dim = 2 * L + 1;
  DecideDesc dd{};
  dd.cur = d_cur;
  dd.ref = refpad;
  dd.cur_stride = W;
  dd.ref_stride = ref_stride;
  dd.W = W;
  dd.H = H;
  dd.mbw = mbw;
  dd.mbh = mbh;
  dd.R = cfg->range;
  dd.L = L;
  dd.metric = cfg->metric;
  dd.filter = 0;
  dd.subpel = cfg->subpel;
  dd.pred = 0;
  dd.bestvec = cfg->best_vector;
  dd.mb_base = 0;
  dd.table = nullptr;
  dd.table_dim = 2 * L + 1;
  dd.mv_out = d_mv;
  dd.dist_out = d_dist;
  dd.int_out = d_int;
  const PadDesc pd{d_ref, refpad, W, H, W, ref_stride, 0};

  StreamDesc* p_sd = (StreamDesc*)base;
  DecideDesc* p_dd = (DecideDesc*)(base + o_dec);
  PadDesc* p_pd = (PadDesc*)(base + o_pad);
  SearchTask* p_task = (SearchTask*)(base + o_task);
  hipLaunchKernelGGL(setup_kernel, dim3((ntasks + 255) / 256), dim3(256), 0, st, sd, dd, pd, p_sd, p_dd, p_pd,
                     p_task, ntasks, mbw, L0 ? 0 : t.G);
  launch_pad_planes(p_pd, 1, pad_plane_workgroups(W, H), st);
  if (L0) {
    launch_wave_search(p_sd, p_task, ntasks, L0, p_dd, true, wave_lean(cfg->subpel, cfg->best_vector), st);
  } else {
    const int stage = (search_stage_dwords(t.B, t.V, t.Q, t.G) + 63) & ~63;
    launch_search(p_sd, p_task, ntasks, cfg->grid_mode ? cfg->grid_mode - 1 : 0, cfg->metric, t.V, false, stage,
                  nullptr, nullptr, p_dd, true, st);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return icu_fail(ICU_E_DEVICE, std::string("launch: ") + hipGetErrorString(e));
  return ICU_OK;
}
