// C ABI over the HIP kernels: estimator handles, batched estimation with
// graph capture, the tile tuner, transforms and a minimal device runtime.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <functional>
#include <queue>
#include <vector>

#include "../../include/icu_hip.h"
#include "icu_device.h"
#include "icu_internal.h"
#include "icu_tx.h"

using namespace icu;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace

int icu::icu_fail(int code, const std::string& msg) { return fail(code, msg); }

namespace {

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(ICU_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_));        \
  } while (0)

}  // namespace

namespace icu {

// ------------------------------------------------------------------ tiling
int quads_for(int L) { return (L - tile_dx0(L)) / 4 + 1; }

// Lane tiles: rows V per lane so that B*V covers the 2L+1 window rows with
// the least waste; MBs per task G so that G*Q*B lanes fill whole 256-lane
// passes and the staged window fits in LDS.  SAD argmin tiles (no cost
// table) also need lane_keys_fit; V = 8 always fits.  Returns V = 0 when a
// requested V is not allowed.
Tiling plan_tiling(int mbw, int L, int metric, bool argmin, int V_req, int G_req) {
  Tiling t{};
  const int N = 2 * L + 1;
  t.Q = quads_for(L);
  static const int kV[4] = {8, 11, 13, 16};
  auto allowed = [&](int v) {
    if (metric == ICU_METRIC_SSD) return v == 8;
    if (argmin) return v != 16 && lane_keys_fit(L, v);
    return true;
  };
  if (V_req && !allowed(V_req)) return t;
  int bestV = 8, bestWaste = 1 << 30;
  for (int v : kV) {
    if (!allowed(v)) continue;
    if (V_req && v != V_req) continue;
    const int waste = ((N + v - 1) / v) * v - N;
    if (waste < bestWaste || (waste == bestWaste && v > bestV)) {
      bestWaste = waste;
      bestV = v;
    }
  }
  t.V = bestV;
  t.B = (N + t.V - 1) / t.V;
  const int U = t.Q * t.B;
  const int gmax = std::min(search_max_group(), mbw);
  int bestG = 1;
  double bestEff = -1;
  for (int g = 1; g <= gmax; g++) {
    if (search_stage_dwords(t.B, t.V, t.Q, g) > search_max_window_dwords()) break;
    if (G_req && g != G_req) continue;
    const int lanes = g * U;
    const double eff = (double)lanes / (256.0 * ((lanes + 255) / 256));
    // prefer full passes; among equals prefer at most 4 passes per task
    const double score = eff - (lanes > 1024 ? 0.01 * (lanes / 1024) : 0.0);
    if (score > bestEff + 1e-9) {
      bestEff = score;
      bestG = g;
    }
  }
  t.G = bestG;
  return t;
}

bool valid_cfg(const icu_me_config& c, std::string* why) {
  if (c.width <= 0 || c.height <= 0 || (c.width & 15) || (c.height & 15)) {
    *why = "width and height must be positive multiples of 16";
    return false;
  }
  if (c.range < 1 || c.range > kMaxRange) {
    *why = "range must be 1..64";
    return false;
  }
  if ((c.metric != 0 && c.metric != 1) || (c.window != 0 && c.window != 1) ||
      (c.filter != 0 && c.filter != 1) || c.subpel < 0 || c.subpel > 2 ||
      (c.predictor != 0 && c.predictor != 1) || (c.best_vector != 0 && c.best_vector != 1)) {
    *why = "mode switch out of range";
    return false;
  }
  if (c.tile_rows && c.tile_rows != 8 && c.tile_rows != 11 && c.tile_rows != 13 && c.tile_rows != 16) {
    *why = "tile_rows must be 0, 8, 11, 13 or 16";
    return false;
  }
  if (c.metric == ICU_METRIC_SSD && c.tile_rows && c.tile_rows != 8) {
    *why = "SSD tiles are 8 rows";
    return false;
  }
  if (c.mbs_per_task < 0 || c.mbs_per_task > search_max_group()) {
    *why = "mbs_per_task must be 0..16";
    return false;
  }
  if (c.grid_mode < 0 || c.grid_mode > 7) {
    *why = "grid_mode must be 0..7";
    return false;
  }
  if (2 * c.range + 1 > 129) {
    *why = "range too large";
    return false;
  }
  return true;
}

}  // namespace icu

namespace {

// ------------------------------------------------------------------ engine
struct StreamPlan {
  icu_me_config cfg;
  Tiling t;
  int L, mbw, mbh, nmb, mb_base;
  bool table;
  long long table_base;
  uint8_t* refpad_base = nullptr;
  uint8_t* refpad = nullptr;
  int ref_stride = 0;
};

struct Engine {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t side = nullptr;  // grid mode 5: window classes 16 and 8 run here, beside the others
  std::vector<StreamPlan> sp;
  int total_mb = 0;
  long long table_total = 0;
  std::vector<SearchTask> h_tasks;   // workgroup tasks (me_search_kernel)
  std::vector<SearchTask> h_wtasks;  // one MB per wave, after h_tasks on the device: grouped by window
                                     // class (kWaveClasses[k] has wave_n[k] tasks) for per-class launches
                                     // of me_wave_kernel, or wave by wave for me_queue_kernel (wave b
                                     // runs [h_woff[b], h_woff[b + 1]))
  static constexpr int kWaveClasses[4] = {64, 32, 16, 8};
  static constexpr int kMaxQueueWaves = 8;
  int wave_n[4] = {0, 0, 0, 0};
  std::vector<int> h_woff;
  int* d_woff = nullptr;             // kMaxQueueWaves x 4 x cus + 1 ints
  bool wave_on = false;              // the wave kernels run the streams they take
  bool lean = false;                 // every wave stream is of the lean modes (wave_lean)
  int queue_waves = 6;               // grid of the queue kernel: waves per SIMD x 4 x CUs
  std::vector<DecideItem> h_items;
  // Descriptor slots.  kEager slots serve eager calls: a call uses the slot
  // already bound to its buffers, or refills the least recently used one
  // after a device synchronisation (its launches, on whatever streams, and
  // its last upload are then done) and uploads it on the call's stream from
  // a pinned mirror.  A call on a bound slot records nothing: an event
  // recorded after every estimate held the next launch on the stream back by
  // 4-5 us.  Slot kGraph belongs to a captured graph, so eager calls never
  // redirect a graph's replays.
  static constexpr int kEager = 3, kGraph = kEager;
  struct DescSet {
    void* d_block = nullptr;  // StreamDesc[n] | DecideDesc[n] | PadDesc[n]
    void* h_block = nullptr;  // pinned host mirror of d_block
    StreamDesc* d_streams = nullptr;
    DecideDesc* d_decide = nullptr;
    PadDesc* d_pad = nullptr;
    std::vector<const uint8_t*> cur, ref;
    int16_t* mv = nullptr;
    int32_t* dist = nullptr;
    int32_t* integ = nullptr;
    bool valid = false;
    bool used = false;        // launched or uploaded since the last device synchronisation
    hipEvent_t fork = nullptr, join = nullptr;  // grid mode 5: the side stream's start and end
    unsigned long long stamp = 0;
  } ds[kEager + 1];
  unsigned long long use_clock = 0;
  size_t desc_bytes = 0;
  long long dev_bytes = 0;  // device memory this engine allocated
  SearchTask* d_tasks = nullptr;
  DecideItem* d_items = nullptr;
  IntResult* d_ires = nullptr;
  uint32_t* d_table = nullptr;
  int* d_wave = nullptr;  // per stream: 2 words (row ticket, error flag)
  uint8_t* d_pf = nullptr;  // predictor-free passes of the median-predictor streams' MBs
  std::vector<int> wave_off;
  std::vector<int> pad_base;  // first padding workgroup of each stream
  int pad_total = 0;          // padding workgroups of all streams
  int stage_dwords = 0;   // one LDS staging buffer of the search kernel
  int cus = 256;          // compute units of the device
  // search grid mode (the tuner may switch it): 0 one task per workgroup
  // (one LDS buffer, so the most workgroups fit); 1 persistent workgroups, as
  // many as stay resident, run every grid-th task with the next task's window
  // streaming into a second buffer; 2 the same with the grid reduced so every
  // workgroup runs equally many tasks; 3 one MB per wave for the streams
  // whose window the wave kernel takes (wave_class), one launch per window
  // class, mode 0 for the rest; 4 the same streams in one persistent launch
  // over per-XCD macroblock queues (me_queue_kernel), mode 0 for the rest;
  // 5 as 3 with the +-8 and +-16 classes on a second stream, so their
  // latency-bound waves run beside the +-32 / +-64 launches (measured
  // better than a stream per class)
  int grid_mode = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;

  ~Engine() { release(); }

  void release() {
    if (gexec) hipGraphExecDestroy(gexec);
    if (graph) hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    for (auto& s : sp)
      if (s.refpad_base) hipFree(s.refpad_base);
    sp.clear();
    bool any_used = false;
    for (auto& d : ds) any_used |= d.used;
    if (any_used) hipDeviceSynchronize();
    for (auto& d : ds) {
      hipFree(d.d_block);
      hipHostFree(d.h_block);
      if (d.fork) hipEventDestroy(d.fork);
      if (d.join) hipEventDestroy(d.join);
      d = DescSet{};
    }
    hipFree(d_tasks);
    hipFree(d_items);
    hipFree(d_ires);
    hipFree(d_table);
    hipFree(d_wave);
    hipFree(d_pf);
    hipFree(d_woff);
    d_woff = nullptr;
    d_tasks = nullptr; d_items = nullptr;
    d_ires = nullptr; d_table = nullptr; d_wave = nullptr; d_pf = nullptr;
    if (own) hipStreamDestroy(own);
    if (side) hipStreamDestroy(side);
    own = nullptr;
    side = nullptr;
  }

  // The wave kernel takes argmin (no cost table) SAD streams whose window
  // half-width has a class (wave_class).
  static bool wave_fits(const StreamPlan& s) {
    return !s.table && s.cfg.metric == ICU_METRIC_SAD && wave_class(s.L) != 0;
  }
  bool any_wave_stream() const {
    for (const auto& s : sp)
      if (wave_fits(s)) return true;
    return false;
  }
  bool wave_stream(const StreamPlan& s) const { return wave_on && wave_fits(s); }

  void build_tasks() {
    h_tasks.clear();
    h_wtasks.clear();
    wave_on = grid_mode >= 3;
    lean = true;
    for (const auto& s : sp)
      if (wave_stream(s)) lean &= wave_lean(s.cfg.subpel, s.cfg.best_vector);
    for (int k = 0; k < 4; k++) {
      wave_n[k] = 0;
      for (size_t si = 0; si < sp.size(); si++) {
        const StreamPlan& s = sp[si];
        if (!wave_stream(s) || wave_class(s.L) != kWaveClasses[k]) continue;
        for (int mb = 0; mb < s.nmb; mb++) h_wtasks.push_back(SearchTask{(int)si, mb, 1});
        wave_n[k] += s.nmb;
      }
    }
    h_woff.clear();
    if (grid_mode == 4) build_queue(4);
    if (grid_mode == 6) build_queue(2);
    for (size_t si = 0; si < sp.size(); si++) {
      const StreamPlan& s = sp[si];
      if (wave_stream(s)) continue;
      for (int row = 0; row < s.mbh; row++)
        for (int c = 0; c < s.mbw; c += s.t.G)
          h_tasks.push_back(SearchTask{(int)si, row * s.mbw + c, std::min(s.t.G, s.mbw - c)});
    }
    // Longest tasks first so a persistent grid ends evenly.
    std::stable_sort(h_tasks.begin(), h_tasks.end(), [&](const SearchTask& a, const SearchTask& b) {
      const StreamPlan &sa = sp[a.stream], &sb = sp[b.stream];
      const long ca = (long)a.count * sa.t.Q * sa.t.B * (sa.t.V + 15);
      const long cb = (long)b.count * sb.t.Q * sb.t.B * (sb.t.V + 15);
      return ca > cb;
    });
    // one staging buffer fits the largest task; the kernel holds two
    stage_dwords = 0;
    for (const auto& s : sp)
      if (!wave_stream(s)) stage_dwords = std::max(stage_dwords, search_stage_dwords(s.t.B, s.t.V, s.t.Q, s.t.G));
    stage_dwords = (stage_dwords + 63) & ~63;
  }

  // The queue kernel's schedule: each window class split into 8 contiguous
  // runs of MBs, run x to XCD x (so an XCD's L2 serves neighbouring windows);
  // an XCD's MBs, costliest first, each to the XCD's wave with the least
  // estimated work so far (workgroup b runs on XCD b mod 8); then the tasks
  // wave by wave.
  // Classes [0, kq) of h_wtasks (grouped by class) become the queue; the
  // rest stay grouped by class after it.
  void build_queue(int kq) {
    const int nw = queue_waves * 4 * cus;
    std::vector<std::vector<SearchTask>> lists(nw);
    int nq_tasks = 0;
    for (int k = 0; k < kq; k++) nq_tasks += wave_n[k];
    for (int x = 0; x < 8; x++) {
      std::vector<std::pair<long, SearchTask>> share;
      for (int k = 0, off = 0; k < kq; off += wave_n[k], k++) {
        const int nq = wave_n[k] / 8, rq = wave_n[k] % 8;
        const int a = off + x * nq + std::min(x, rq), len = nq + (x < rq ? 1 : 0);
        for (int i = a; i < a + len; i++) {
          const SearchTask& t = h_wtasks[i];
          const StreamPlan& s = sp[t.stream];
          share.push_back({wave_task_cost(16 * (t.mb % s.mbw), 16 * (t.mb / s.mbw), s.cfg.width, s.cfg.height, s.L),
                           t});
        }
      }
      std::stable_sort(share.begin(), share.end(),
                       [](const std::pair<long, SearchTask>& a, const std::pair<long, SearchTask>& b) {
                         return a.first > b.first;
                       });
      // least-loaded wave first; ties to the lower wave index
      typedef std::pair<long, int> Load;
      std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
      for (int b = x; b < nw; b += 8) heap.push({0, b});
      if (heap.empty()) continue;
      for (const auto& ct : share) {
        Load l = heap.top();
        heap.pop();
        lists[l.second].push_back(ct.second);
        heap.push({l.first + ct.first, l.second});
      }
    }
    std::vector<SearchTask> out;
    out.reserve(h_wtasks.size());
    h_woff.assign(nw + 1, 0);
    for (int b = 0; b < nw; b++) {
      h_woff[b] = (int)out.size();
      out.insert(out.end(), lists[b].begin(), lists[b].end());
    }
    h_woff[nw] = (int)out.size();
    out.insert(out.end(), h_wtasks.begin() + nq_tasks, h_wtasks.end());
    h_wtasks.swap(out);
  }

  int init(const icu_me_config* cfgs, int n) {
    HIP_TRY(hipGetDevice(&device));
    // a blocking stream: work on it is ordered with the null stream, so a
    // caller's plain hipMemcpy after an asynchronous call sees the results
    HIP_TRY(hipStreamCreate(&own));
    HIP_TRY(hipStreamCreate(&side));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    cus = prop.multiProcessorCount;
    if (cfgs[0].grid_mode) grid_mode = cfgs[0].grid_mode - 1;
    sp.resize(n);
    total_mb = 0;
    table_total = 0;
    int wave_total = 0;
    for (int i = 0; i < n; i++) {
      StreamPlan& s = sp[i];
      s.cfg = cfgs[i];
      s.L = s.cfg.window == ICU_WINDOW_FULL ? s.cfg.range : s.cfg.range - 1;
      s.mbw = s.cfg.width / 16;
      s.mbh = s.cfg.height / 16;
      s.nmb = s.mbw * s.mbh;
      s.mb_base = total_mb;
      total_mb += s.nmb;
      s.table = s.cfg.filter || s.cfg.predictor;
      s.t = plan_tiling(s.mbw, s.L, s.cfg.metric, !s.table, s.cfg.tile_rows, s.cfg.mbs_per_task);
      if (s.t.V == 0)
        return fail(ICU_E_ARG, "tile_rows " + std::to_string(s.cfg.tile_rows) +
                                   " is too tall for exact SAD argmin keys at range " +
                                   std::to_string(s.cfg.range));
      s.table_base = table_total;
      if (s.table) table_total += (long long)s.nmb * (2 * s.L + 1) * (2 * s.L + 1);
      s.ref_stride = ((s.cfg.width + 2 * kPad) + 255) & ~255;
      const size_t bytes = (size_t)s.ref_stride * (s.cfg.height + 2 * kPad);
      HIP_TRY(hipMalloc(&s.refpad_base, bytes));
      dev_bytes += bytes;
      s.refpad = s.refpad_base + (size_t)kPad * s.ref_stride + kPad;
      wave_off.push_back(wave_total);
      wave_total += 2;
      pad_base.push_back(pad_total);
      pad_total += pad_plane_workgroups(s.cfg.width, s.cfg.height);
    }
    // unless a grid mode is asked for, streams the wave kernels take run one
    // MB per wave (measured faster than the workgroup tiles at +-31 / +-32),
    // one launch per window class; a batch's +-8 / +-16 classes on a second
    // stream (the fastest untuned plan on the benchmark's mixed batch)
    if (!cfgs[0].grid_mode && any_wave_stream()) grid_mode = n > 1 ? 5 : 3;
    build_tasks();
    // descriptor block: three arrays, each starting on a 256-byte boundary
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_dec = up(sizeof(StreamDesc) * n), o_pad = o_dec + up(sizeof(DecideDesc) * n);
    desc_bytes = o_pad + up(sizeof(PadDesc) * n);
    for (auto& d : ds) {
      HIP_TRY(hipMalloc(&d.d_block, desc_bytes));
      dev_bytes += desc_bytes;
      HIP_TRY(hipHostMalloc(&d.h_block, desc_bytes, hipHostMallocDefault));
      // fork / join order two device streams: a device-scope release is
      // enough (a system-scope one writes the caches back)
      HIP_TRY(hipEventCreateWithFlags(&d.fork, hipEventDisableTiming | hipEventReleaseToDevice));
      HIP_TRY(hipEventCreateWithFlags(&d.join, hipEventDisableTiming | hipEventReleaseToDevice));
      d.d_streams = (StreamDesc*)d.d_block;
      d.d_decide = (DecideDesc*)((char*)d.d_block + o_dec);
      d.d_pad = (PadDesc*)((char*)d.d_block + o_pad);
      d.cur.assign(n, nullptr);
      d.ref.assign(n, nullptr);
    }
    // every task holds at least one MB: total_mb entries fit any retiling
    HIP_TRY(hipMalloc(&d_tasks, sizeof(SearchTask) * total_mb));
    HIP_TRY(hipMalloc(&d_ires, sizeof(IntResult) * total_mb));
    HIP_TRY(hipMalloc(&d_wave, sizeof(int) * wave_total));
    HIP_TRY(hipMalloc(&d_woff, sizeof(int) * (kMaxQueueWaves * 4 * cus + 1)));
    dev_bytes += sizeof(int) * (kMaxQueueWaves * 4 * cus + 1);
    dev_bytes += (sizeof(SearchTask) + sizeof(IntResult)) * total_mb + sizeof(int) * wave_total;
    HIP_TRY(hipMemset(d_wave, 0, sizeof(int) * wave_total));
    if (table_total) HIP_TRY(hipMalloc(&d_table, sizeof(uint32_t) * table_total));
    dev_bytes += sizeof(uint32_t) * table_total;
    bool any_pred = false;
    for (const auto& s : sp) any_pred |= s.cfg.predictor != 0;
    if (any_pred) {
      HIP_TRY(hipMalloc(&d_pf, (size_t)predfree_bytes() * total_mb));
      dev_bytes += (long long)predfree_bytes() * total_mb;
    }
    HIP_TRY(upload_tasks());
    h_items.clear();
    for (int i = 0; i < n; i++)
      if (!sp[i].cfg.predictor)
        for (int mb = 0; mb < sp[i].nmb; mb++) h_items.push_back(DecideItem{i, mb});
    if (!h_items.empty()) {
      HIP_TRY(hipMalloc(&d_items, sizeof(DecideItem) * h_items.size()));
      dev_bytes += sizeof(DecideItem) * h_items.size();
      HIP_TRY(hipMemcpy(d_items, h_items.data(), sizeof(DecideItem) * h_items.size(), hipMemcpyHostToDevice));
    }
    return ICU_OK;
  }

  bool bound_to(const DescSet& D, const uint8_t* const* curs, const uint8_t* const* refs, int16_t* mv,
                int32_t* dist, int32_t* integ) const {
    bool same = D.valid && mv == D.mv && dist == D.dist && integ == D.integ;
    for (size_t i = 0; same && i < sp.size(); i++) same = D.cur[i] == curs[i] && D.ref[i] == refs[i];
    return same;
  }

  // Slot for an eager call on stream st: the one bound to these buffers, or
  // the least recently used eager slot refilled (after its last launch) and
  // uploaded on st ahead of the launches that read it.
  int bind_eager(const uint8_t* const* curs, const uint8_t* const* refs, int16_t* mv, int32_t* dist,
                 int32_t* integ, hipStream_t st, int* slot) {
    int pick_slot = 0;
    for (int k = 0; k < kEager; k++) {
      if (bound_to(ds[k], curs, refs, mv, dist, integ)) {
        ds[k].stamp = ++use_clock;
        *slot = k;
        return ICU_OK;
      }
      if (ds[k].stamp < ds[pick_slot].stamp) pick_slot = k;
    }
    DescSet& D = ds[pick_slot];
    if (D.used) {
      HIP_TRY(hipDeviceSynchronize());  // its launches and its last upload are done
      for (auto& x : ds) x.used = false;
    }
    int rc = fill(D, curs, refs, mv, dist, integ);
    if (rc) return rc;
    // the mirror stays in use until the copy ran
    HIP_TRY(hipMemcpyAsync(D.d_block, D.h_block, desc_bytes, hipMemcpyHostToDevice, st));
    D.used = true;
    D.stamp = ++use_clock;
    *slot = pick_slot;
    return ICU_OK;
  }

  // The graph's slot, written synchronously before capture (replays read it).
  int bind_graph(const uint8_t* const* curs, const uint8_t* const* refs, int16_t* mv, int32_t* dist) {
    DescSet& D = ds[kGraph];
    if (bound_to(D, curs, refs, mv, dist, nullptr)) return ICU_OK;
    HIP_TRY(hipDeviceSynchronize());  // earlier replays may still read it
    int rc = fill(D, curs, refs, mv, dist, nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(D.d_block, D.h_block, desc_bytes, hipMemcpyHostToDevice));
    return ICU_OK;
  }

  // Write a slot's descriptors into its pinned host mirror.
  int fill(DescSet& D, const uint8_t* const* curs, const uint8_t* const* refs, int16_t* mv, int32_t* dist,
           int32_t* integ) {
    for (size_t i = 0; i < sp.size(); i++)
      if (((uintptr_t)curs[i] & 15) != 0)
        return fail(ICU_E_ARG, "device current planes must be 16-byte aligned");
    const size_t o_dec = (char*)D.d_decide - (char*)D.d_block, o_pad = (char*)D.d_pad - (char*)D.d_block;
    StreamDesc* h_streams = (StreamDesc*)D.h_block;
    DecideDesc* h_decide = (DecideDesc*)((char*)D.h_block + o_dec);
    PadDesc* h_pad = (PadDesc*)((char*)D.h_block + o_pad);
    for (size_t i = 0; i < sp.size(); i++) {
      const StreamPlan& s = sp[i];
      h_pad[i] = PadDesc{refs[i], s.refpad, s.cfg.width, s.cfg.height, s.cfg.width, s.ref_stride, pad_base[i]};
    }
    for (size_t i = 0; i < sp.size(); i++) {
      const StreamPlan& s = sp[i];
      StreamDesc& d = h_streams[i];
      d.cur = curs[i];
      d.ref = s.refpad;
      d.cur_stride = s.cfg.width;
      d.ref_stride = s.ref_stride;
      d.W = s.cfg.width;
      d.H = s.cfg.height;
      d.mbw = s.mbw;
      d.mbh = s.mbh;
      d.R = s.cfg.range;
      d.L = s.L;
      d.V = s.t.V;
      d.B = s.t.B;
      d.Q = s.t.Q;
      d.G = s.t.G;
      d.metric = s.cfg.metric;
      d.mb_base = s.mb_base;
      d.table_base_valid = s.table;
      d.table_base = s.table_base;
      d.table_dim = 2 * s.L + 1;
      DecideDesc& e = h_decide[i];
      e.cur = curs[i];
      e.ref = s.refpad;
      e.cur_stride = s.cfg.width;
      e.ref_stride = s.ref_stride;
      e.W = s.cfg.width;
      e.H = s.cfg.height;
      e.mbw = s.mbw;
      e.mbh = s.mbh;
      e.R = s.cfg.range;
      e.L = s.L;
      e.metric = s.cfg.metric;
      e.filter = s.cfg.filter;
      e.subpel = s.cfg.subpel;
      e.pred = s.cfg.predictor;
      e.bestvec = s.cfg.best_vector;
      e.mb_base = s.mb_base;
      e.table = s.table ? d_table : nullptr;
      e.table_base = s.table_base;
      e.table_dim = 2 * s.L + 1;
      e.ires = d_ires;
      e.mv_out = mv;
      e.dist_out = dist;
      e.int_out = integ;
      e.ticket = d_wave + wave_off[i];
    }
    for (size_t i = 0; i < sp.size(); i++) {
      D.cur[i] = curs[i];
      D.ref[i] = refs[i];
    }
    D.mv = mv;
    D.dist = dist;
    D.integ = integ;
    D.valid = true;
    return ICU_OK;
  }

  // Enqueue the whole estimate on `st` (buffers already bound).
  // ev (nullable): 4 events recorded before padding, search, decision, end.
  // search_reps > 1 (timing only): the search launch is repeated; it reads
  // only the padded references and the current planes, so every repeat
  // writes the same results.
  int decide_launches = 0;  // decision kernels of the last enqueue
  int enqueue(int set, hipStream_t st, hipEvent_t* ev = nullptr, int search_reps = 1) {
    StreamDesc* d_streams = ds[set].d_streams;
    DecideDesc* d_decide = ds[set].d_decide;
    if (ev) HIP_TRY(hipEventRecord(ev[0], st));
    launch_pad_planes(ds[set].d_pad, (int)sp.size(), pad_total, st);
    if (ev) HIP_TRY(hipEventRecord(ev[1], st));
    int vfix = -1;
    bool any_table = false;
    for (const auto& x : sp) {
      if (wave_stream(x)) continue;
      vfix = vfix < 0 || vfix == x.t.V ? x.t.V : 0;
      any_table |= x.table;
    }
    for (int r = 0; r < search_reps; r++) {
      if (!h_tasks.empty())
        launch_search(d_streams, d_tasks, (int)h_tasks.size(), grid_mode >= 3 ? 0 : grid_mode, sp[0].cfg.metric,
                      std::max(vfix, 0), any_table, stage_dwords, d_ires, d_table, d_decide, sp.size() == 1, st);
      if (grid_mode == 4) {
        launch_wave_queue(d_streams, d_tasks + h_tasks.size(), d_woff, queue_waves * 4 * cus, false, lean, d_decide,
                          st);
      } else if (grid_mode == 6) {
        // classes 64 and 32 in one persistent launch on st, 16 and 8 one
        // launch each on the side stream
        const int nq = wave_n[0] + wave_n[1];
        const bool split = nq > 0 && (wave_n[2] + wave_n[3]) > 0;
        if (split) {
          HIP_TRY(hipEventRecord(ds[set].fork, st));
          HIP_TRY(hipStreamWaitEvent(side, ds[set].fork, 0));
        }
        if (nq)
          launch_wave_queue(d_streams, d_tasks + h_tasks.size(), d_woff, queue_waves * 4 * cus, true, lean, d_decide,
                            st);
        for (int k = 2, off = (int)h_tasks.size() + nq; k < 4; off += wave_n[k], k++)
          launch_wave_search(d_streams, d_tasks + off, wave_n[k], kWaveClasses[k], d_decide, true, lean,
                             split ? side : st);
        if (split) {
          HIP_TRY(hipEventRecord(ds[set].join, side));
          HIP_TRY(hipStreamWaitEvent(st, ds[set].join, 0));
        }
      } else {
        // mode 5: classes 64 and 32 on st, 16 and 8 on the side stream,
        // which starts after the padding and which st waits for at the end
        const bool split = grid_mode == 5 && (wave_n[0] + wave_n[1]) > 0 && (wave_n[2] + wave_n[3]) > 0;
        if (split) {
          HIP_TRY(hipEventRecord(ds[set].fork, st));
          HIP_TRY(hipStreamWaitEvent(side, ds[set].fork, 0));
        }
        for (int k = 0, off = (int)h_tasks.size(); k < 4; off += wave_n[k], k++)
          launch_wave_search(d_streams, d_tasks + off, wave_n[k], kWaveClasses[k], d_decide, true, lean,
                             split && k >= 2 ? side : st);
        if (split) {
          HIP_TRY(hipEventRecord(ds[set].join, side));
          HIP_TRY(hipStreamWaitEvent(st, ds[set].join, 0));
        }
      }
    }
    if (ev) HIP_TRY(hipEventRecord(ev[2], st));
    // without cost tables the search kernel decides every MB itself
    decide_launches = 0;
    if (any_table && !h_items.empty()) {
      launch_decide_parallel(d_decide, d_items, (int)h_items.size(), st);
      decide_launches++;
    }
    for (size_t i = 0; i < sp.size(); i++) {
      if (!sp[i].cfg.predictor) continue;
      decide_launches += 2;
      // the predictor-free pass resets the row ticket and marks every
      // vector pending; the error flag stays set until check_wave_errors
      // reads it
      uint8_t* pf = d_pf + (size_t)predfree_bytes() * sp[i].mb_base;
      launch_decide_predfree(d_decide + i, pf, sp[i].nmb, st);
      launch_decide_wavefront(d_decide + i, sp[i].table ? 2 * sp[i].L + 1 : 0, sp[i].mbh, pf, st);
    }
    if (ev) HIP_TRY(hipEventRecord(ev[3], st));
    HIP_TRY(hipGetLastError());
    if (set != kGraph) ds[set].used = true;
    return ICU_OK;
  }

  // Reads (and clears) the median-predictor wavefronts' error flags; the
  // device work that may set them must have finished.
  // Workgroup tasks, then wave tasks, into d_tasks (total_mb entries hold
  // both: every task covers at least one MB).
  hipError_t upload_tasks() {
    hipError_t e = hipSuccess;
    if (!h_tasks.empty())
      e = hipMemcpy(d_tasks, h_tasks.data(), sizeof(SearchTask) * h_tasks.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !h_wtasks.empty())
      e = hipMemcpy(d_tasks + h_tasks.size(), h_wtasks.data(), sizeof(SearchTask) * h_wtasks.size(),
                    hipMemcpyHostToDevice);
    if (e == hipSuccess && !h_woff.empty())
      e = hipMemcpy(d_woff, h_woff.data(), sizeof(int) * h_woff.size(), hipMemcpyHostToDevice);
    return e;
  }

  // Only the handle's own stream is used (a null-stream copy would wait for
  // every other handle's work too).
  int check_wave_errors() {
    bool timed_out = false;
    for (size_t i = 0; i < sp.size(); i++) {
      if (!sp[i].cfg.predictor) continue;
      int* flag = d_wave + wave_off[i] + 1;
      int v = 0;
      HIP_TRY(hipMemcpyAsync(&v, flag, sizeof(v), hipMemcpyDeviceToHost, own));
      HIP_TRY(hipStreamSynchronize(own));
      if (v) {
        timed_out = true;
        launch_fill_i32(flag, 0, 1, own);
        HIP_TRY(hipStreamSynchronize(own));
      }
    }
    return timed_out ? fail(ICU_E_STATE, "median-predictor wavefront timed out") : ICU_OK;
  }

  int retile(int si, Tiling t) {
    sp[si].t = t;
    return replan();
  }

  // Rebuilds and uploads the task lists after a change of tiling or grid mode.
  int replan() {
    build_tasks();
    if ((int)(h_tasks.size() + h_wtasks.size()) > total_mb) return fail(ICU_E_STATE, "task list exceeds its allocation");
    // kernels of an earlier call may still read the task list
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(upload_tasks());
    for (auto& d : ds) d.valid = false;
    return ICU_OK;
  }
};

}  // namespace

// ---------------------------------------------------------------- handles
struct icu_me {
  Engine eng;
  // staging for the host entry points
  uint8_t* d_cur = nullptr;
  uint8_t* d_ref = nullptr;
  int16_t* d_mv = nullptr;
  int32_t* d_dist = nullptr;
  std::vector<uint8_t> h8;
  ~icu_me() {
    hipFree(d_cur);
    hipFree(d_ref);
    hipFree(d_mv);
    hipFree(d_dist);
  }
};

struct icu_me_batch {
  Engine eng;
};

int* icu::me_wave_flag(icu_me* h) {
  const Engine& e = h->eng;
  if (!e.sp[0].cfg.predictor) return nullptr;
  return e.d_wave + e.wave_off[0] + 1;
}

int icu::me_check(icu_me* h) { return h->eng.check_wave_errors(); }

long long icu::me_device_bytes(const icu_me* h) {
  const long long px = (long long)h->eng.sp[0].cfg.width * h->eng.sp[0].cfg.height;
  return h->eng.dev_bytes + 2 * px + 8 * (px / 256);
}

static hipStream_t pick(void* stream, Engine& e) { return stream ? (hipStream_t)stream : e.own; }

extern "C" {

const char* icu_last_error(void) { return g_err.c_str(); }

int icu_me_config_reference(icu_me_config* cfg, int width, int height, int motion_range_qpel) {
  if (!cfg) return fail(ICU_E_ARG, "null config");
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->width = width;
  cfg->height = height;
  cfg->range = motion_range_qpel / 4;
  cfg->metric = ICU_METRIC_SSD;
  cfg->window = ICU_WINDOW_REFERENCE;
  cfg->filter = 1;
  cfg->subpel = 2;
  cfg->predictor = 1;
  cfg->best_vector = 1;
  return ICU_OK;
}

int icu_me_create(const icu_me_config* cfg, icu_me** out) {
  if (!cfg || !out) return fail(ICU_E_ARG, "null argument");
  std::string why;
  if (!valid_cfg(*cfg, &why)) return fail(ICU_E_ARG, why);
  icu_me* h = new icu_me();
  int rc = h->eng.init(cfg, 1);
  const size_t px = (size_t)cfg->width * cfg->height, nmb = px / 256;
  if (rc == ICU_OK) {
    if (hipMalloc(&h->d_cur, px) != hipSuccess || hipMalloc(&h->d_ref, px) != hipSuccess ||
        hipMalloc(&h->d_mv, nmb * 4) != hipSuccess || hipMalloc(&h->d_dist, nmb * 4) != hipSuccess)
      rc = fail(ICU_E_DEVICE, "device allocation failed");
  }
  if (rc != ICU_OK) {
    delete h;
    return rc;
  }
  *out = h;
  return ICU_OK;
}

int icu_me_destroy(icu_me* h) {
  delete h;
  return ICU_OK;
}

static int to_u8(const int16_t* src, size_t n, std::vector<uint8_t>& dst, size_t off) {
  int bad = 0;
  for (size_t i = 0; i < n; i++) {
    const int v = src[i];
    bad |= (v < 0) | (v > 255);
    dst[off + i] = (uint8_t)v;
  }
  return bad ? fail(ICU_E_RANGE, "input sample outside 0..255") : ICU_OK;
}

int icu_me_estimate_device(icu_me* h, const uint8_t* d_cur, const uint8_t* d_ref, int16_t* d_mv,
                           int32_t* d_dist, int32_t* d_int, void* stream) {
  if (!h || !d_cur || !d_ref || !d_mv || !d_dist) return fail(ICU_E_ARG, "null argument");
  Engine& e = h->eng;
  const uint8_t* curs[1] = {d_cur};
  const uint8_t* refs[1] = {d_ref};
  const hipStream_t st = pick(stream, e);
  int slot = 0;
  int rc = e.bind_eager(curs, refs, d_mv, d_dist, d_int, st, &slot);
  if (rc) return rc;
  return e.enqueue(slot, st);
}

int icu_me_estimate_timed(icu_me* h, const uint8_t* d_cur, const uint8_t* d_ref, int16_t* d_mv,
                          int32_t* d_dist, void* stream, int search_reps, float* stage_ms) {
  if (!h || !d_cur || !d_ref || !d_mv || !d_dist || !stage_ms) return fail(ICU_E_ARG, "null argument");
  if (search_reps < 1 || search_reps > 1000) return fail(ICU_E_ARG, "search_reps must be 1..1000");
  Engine& e = h->eng;
  const uint8_t* curs[1] = {d_cur};
  const uint8_t* refs[1] = {d_ref};
  const hipStream_t st = pick(stream, e);
  int slot = 0;
  int rc = e.bind_eager(curs, refs, d_mv, d_dist, nullptr, st, &slot);
  if (rc) return rc;
  hipEvent_t ev[4];
  for (auto& x : ev) HIP_TRY(hipEventCreateWithFlags(&x, hipEventReleaseToDevice));
  rc = e.enqueue(slot, st, ev, search_reps);
  if (!rc) {
    HIP_TRY(hipEventSynchronize(ev[3]));
    for (int i = 0; i < 3; i++) HIP_TRY(hipEventElapsedTime(&stage_ms[i], ev[i], ev[i + 1]));
    stage_ms[1] /= (float)search_reps;
    // no decision kernels: the last pair brackets nothing but its own gap
    if (e.decide_launches == 0) stage_ms[2] = 0.f;
  }
  for (auto& x : ev) hipEventDestroy(x);
  return rc;
}

int icu_me_estimate(icu_me* h, const int16_t* cur, const int16_t* ref, const uint8_t* include,
                    int16_t* mv_out, int32_t* dist_out, long* avg_out) {
  if (!h || !cur || !ref || !mv_out) return fail(ICU_E_ARG, "null argument");
  Engine& e = h->eng;
  const icu_me_config& c = e.sp[0].cfg;
  const size_t px = (size_t)c.width * c.height, nmb = px / 256;
  h->h8.resize(2 * px);
  int rc = to_u8(cur, px, h->h8, 0);
  if (!rc) rc = to_u8(ref, px, h->h8, px);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(h->d_cur, h->h8.data(), px, hipMemcpyHostToDevice, e.own));
  HIP_TRY(hipMemcpyAsync(h->d_ref, h->h8.data() + px, px, hipMemcpyHostToDevice, e.own));
  rc = icu_me_estimate_device(h, h->d_cur, h->d_ref, h->d_mv, h->d_dist, nullptr, e.own);
  if (rc) return rc;
  std::vector<int32_t> dist(nmb);
  HIP_TRY(hipMemcpyAsync(mv_out, h->d_mv, nmb * 4, hipMemcpyDeviceToHost, e.own));
  HIP_TRY(hipMemcpyAsync(dist.data(), h->d_dist, nmb * 4, hipMemcpyDeviceToHost, e.own));
  HIP_TRY(hipStreamSynchronize(e.own));
  rc = e.check_wave_errors();
  if (rc) return rc;
  if (dist_out) std::memcpy(dist_out, dist.data(), nmb * 4);
  if (avg_out) {
    long total = 0;
    int inc = 0;
    if (include)
      for (size_t i = 0; i < nmb; i++)
        if (include[i]) {
          total += dist[i];
          inc++;
        }
    *avg_out = inc ? total / inc : 0;
  }
  return ICU_OK;
}

int icu_me_check(icu_me* h) {
  if (!h) return fail(ICU_E_ARG, "null handle");
  HIP_TRY(hipDeviceSynchronize());
  return h->eng.check_wave_errors();
}

int icu_me_tie_count(icu_me* h, int* tie_mbs) {
  if (!h || !tie_mbs) return fail(ICU_E_ARG, "null argument");
  Engine& e = h->eng;
  *tie_mbs = 0;
  if (!e.d_pf) return ICU_OK;
  HIP_TRY(hipDeviceSynchronize());
  const int pfb = predfree_bytes();
  std::vector<int32_t> pf((size_t)pfb / 4 * e.total_mb);
  HIP_TRY(hipMemcpy(pf.data(), e.d_pf, pf.size() * 4, hipMemcpyDeviceToHost));
  // the tie flag is the last int of each record
  for (int i = 0; i < e.total_mb; i++) *tie_mbs += pf[(size_t)(i + 1) * (pfb / 4) - 1] != 0;
  return ICU_OK;
}

int icu_me_debug_set_wave_error(icu_me* h) {
  if (!h) return fail(ICU_E_ARG, "null handle");
  int* flag = icu::me_wave_flag(h);
  if (!flag) return fail(ICU_E_STATE, "the handle has no median-predictor wavefront");
  const int one = 1;
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(flag, &one, sizeof(one), hipMemcpyHostToDevice));
  return ICU_OK;
}

int icu_me_get_tiling(icu_me* h, int* tile_rows, int* mbs_per_task, int* grid_mode) {
  if (!h) return fail(ICU_E_ARG, "null handle");
  if (tile_rows) *tile_rows = h->eng.sp[0].t.V;
  if (mbs_per_task) *mbs_per_task = h->eng.sp[0].t.G;
  if (grid_mode) *grid_mode = h->eng.grid_mode;
  return ICU_OK;
}

namespace {
// Times `reps` estimates of the handle's current tiling on its own stream.
int time_estimates(icu_me* h, const uint8_t* d_cur, const uint8_t* d_ref, int reps, hipEvent_t a,
                   hipEvent_t b, float* ms) {
  Engine& e = h->eng;
  int rc = icu_me_estimate_device(h, d_cur, d_ref, h->d_mv, h->d_dist, nullptr, e.own);  // warm-up
  if (rc) return rc;
  HIP_TRY(hipEventRecord(a, e.own));
  for (int r = 0; r < reps; r++) {
    rc = icu_me_estimate_device(h, d_cur, d_ref, h->d_mv, h->d_dist, nullptr, e.own);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(b, e.own));
  HIP_TRY(hipEventSynchronize(b));
  HIP_TRY(hipEventElapsedTime(ms, a, b));
  return ICU_OK;
}
}  // namespace

int icu_me_autotune(icu_me* h, const uint8_t* d_cur, const uint8_t* d_ref, int reps) {
  if (!h || !d_cur || !d_ref) return fail(ICU_E_ARG, "null argument");
  Engine& e = h->eng;
  StreamPlan& s = e.sp[0];
  if (reps < 1) reps = 1;
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    hipEventDestroy(a);
    return fail(ICU_E_DEVICE, "hipEventCreate failed");
  }
  Tiling best = s.t;
  int best_mode = e.grid_mode;
  float best_ms = 1e30f;
  int rc = ICU_OK;
  static const int kV[4] = {8, 11, 13, 16};
  for (int mode = 0; mode < 3 && rc == ICU_OK; mode++) {
    for (int v : kV) {
      if (rc != ICU_OK) break;
      if (s.cfg.metric == ICU_METRIC_SSD && v != 8) continue;
      if ((2 * s.L + 1 + v - 1) / v * v - (2 * s.L + 1) > v / 2 + 2) continue;  // too much row waste
      for (int g = 1; g <= std::min(search_max_group(), s.mbw); g++) {
        Tiling t = plan_tiling(s.mbw, s.L, s.cfg.metric, !s.table, v, g);
        if (t.V != v || t.G != g) continue;
        const int lanes = g * t.Q * t.B;
        if (lanes < 192 && g < s.mbw) continue;  // leaves most of a workgroup idle
        e.grid_mode = mode;
        float ms = 0;
        if ((rc = e.retile(0, t)) != ICU_OK || (rc = time_estimates(h, d_cur, d_ref, reps, a, b, &ms)) != ICU_OK)
          break;
        if (ms < best_ms) {
          best_ms = ms;
          best = t;
          best_mode = mode;
        }
      }
    }
  }
  // one MB per wave, where the stream's window takes it: one launch per
  // window class, or the queue kernel
  for (int mode = 3; mode <= 4 && rc == ICU_OK && e.any_wave_stream(); mode++) {
    e.grid_mode = mode;
    float ms = 0;
    if ((rc = e.retile(0, s.t)) == ICU_OK && (rc = time_estimates(h, d_cur, d_ref, reps, a, b, &ms)) == ICU_OK &&
        ms < best_ms) {
      best_ms = ms;
      best = s.t;
      best_mode = mode;
    }
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (rc != ICU_OK) return rc;
  e.grid_mode = best_mode;
  return e.retile(0, best);
}

int icu_code_frame_qps_device(icu_me* h, const uint8_t* d_cur, const int16_t* d_mv, int nqp,
                              const int* qps, int intra, int cols_first, int16_t* const* d_levels,
                              uint8_t* const* d_recon, int16_t* d_pred, void* stream) {
  if (!h || !d_cur || !d_mv || !qps || !d_levels || !d_recon) return fail(ICU_E_ARG, "null argument");
  if (nqp < 1 || nqp > kMaxQpSet) return fail(ICU_E_ARG, "nqp must be 1..20");
  QpSet qs{};
  qs.n = nqp;
  // the coder reads and writes whole rows of 4 samples
  if (((uintptr_t)d_cur & 3) || ((uintptr_t)d_pred & 7))
    return fail(ICU_E_ARG, "d_cur must be 4-byte, d_pred 8-byte aligned");
  for (int i = 0; i < nqp; i++) {
    if (qps[i] < 0 || qps[i] > 51) return fail(ICU_E_ARG, "qp must be 0..51");
    if (!d_levels[i] || !d_recon[i]) return fail(ICU_E_ARG, "null output");
    if (((uintptr_t)d_recon[i] & 3) || ((uintptr_t)d_levels[i] & 7))
      return fail(ICU_E_ARG, "d_recon must be 4-byte, d_levels 8-byte aligned");
    qs.qp[i] = qps[i];
    qs.levels[i] = d_levels[i];
    qs.recon[i] = d_recon[i];
  }
  Engine& e = h->eng;
  const StreamPlan& s = e.sp[0];
  launch_code_frame(d_cur, s.cfg.width, s.refpad, s.ref_stride, s.cfg.width, s.cfg.height, d_mv, qs,
                    intra ? 1 : 0, cols_first ? 1 : 0, d_pred, pick(stream, e));
  HIP_TRY(hipGetLastError());
  return ICU_OK;
}

int icu_code_frame_device(icu_me* h, const uint8_t* d_cur, const int16_t* d_mv, int qp, int intra,
                          int cols_first, int16_t* d_levels, uint8_t* d_recon, int16_t* d_pred,
                          void* stream) {
  return icu_code_frame_qps_device(h, d_cur, d_mv, 1, &qp, intra, cols_first, &d_levels, &d_recon,
                                   d_pred, stream);
}

// ------------------------------------------------------------ transforms
static int check_tx(long n, int qp, int mode, int layout) {
  if (n < 0) return fail(ICU_E_ARG, "negative block count");
  if (qp < 0 || qp > 51) return fail(ICU_E_ARG, "qp must be 0..51");
  if (mode < 0 || mode > 2) return fail(ICU_E_ARG, "mode must be 0, 1 or 2");
  if (layout != 0 && layout != 1) return fail(ICU_E_ARG, "layout must be 0 or 1");
  if (layout == 1 && (n % 16)) return fail(ICU_E_ARG, "layout 1 needs whole macroblocks");
  return ICU_OK;
}

int icu_fwd4x4_device(const int16_t* d_in, int16_t* d_out, long n, int qp, int intra, int mode,
                      int layout, void* stream) {
  int rc = check_tx(n, qp, mode, layout);
  if (rc) return rc;
  launch_fwd4x4(d_in, d_out, n, qp, intra ? 1 : 0, mode, layout, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return ICU_OK;
}

int icu_inv4x4_device(const int16_t* d_in, int16_t* d_out, long n, int qp, int mode, int layout,
                      int cols_first, const int* d_w, void* stream) {
  int rc = check_tx(n, qp, mode, layout);
  if (rc) return rc;
  launch_inv4x4(d_in, d_out, n, qp, mode, layout, cols_first ? 1 : 0, d_w, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return ICU_OK;
}

// Host buffers.  Up to kHostTxBlocks blocks (a per-block codec call is
// one) run on the CPU with the kernels' own arithmetic (icu_tx.h): a device
// round trip would cost allocations, two copies and a launch per call.
// Larger batches go through the device kernels on buffers freed on every
// path out.
constexpr long kHostTxBlocks = 16384;

namespace {
struct DeviceScratch {
  void* p = nullptr;
  ~DeviceScratch() {
    if (p) hipFree(p);
  }
};

void tx_cpu(const int16_t* in, int16_t* out, long n, int qp, int intra, int mode, int layout, int inverse,
            int cols_first, const int* w) {
  for (long blk = 0; blk < n; blk++) {
    int16_t b[16];
    for (int p = 0; p < 16; p++) b[p] = in[icu::tx::elem_off(blk, p, layout)];
    if (inverse) icu::tx::inv4x4(b, qp, mode, cols_first, w);
    else icu::tx::fwd4x4(b, qp, intra, mode);
    for (int p = 0; p < 16; p++) out[icu::tx::elem_off(blk, p, layout)] = b[p];
  }
}
}  // namespace

static int tx_host(const int16_t* in, int16_t* out, long n, int qp, int intra, int mode, int layout,
                   int inverse, int cols_first, const int* w) {
  if (!in || !out) return fail(ICU_E_ARG, "null argument");
  int rc = check_tx(n, qp, mode, layout);
  if (rc || n == 0) return rc;
  if (n <= kHostTxBlocks) {
    tx_cpu(in, out, n, qp, intra, mode, layout, inverse, cols_first, w);
    return ICU_OK;
  }
  const size_t bytes = (size_t)n * 16 * sizeof(int16_t);
  DeviceScratch di, dout, dw;
  HIP_TRY(hipMalloc(&di.p, bytes));
  HIP_TRY(hipMalloc(&dout.p, bytes));
  if (w) HIP_TRY(hipMalloc(&dw.p, 16 * sizeof(int)));
  HIP_TRY(hipMemcpy(di.p, in, bytes, hipMemcpyHostToDevice));
  if (w) HIP_TRY(hipMemcpy(dw.p, w, 16 * sizeof(int), hipMemcpyHostToDevice));
  if (inverse)
    launch_inv4x4((const int16_t*)di.p, (int16_t*)dout.p, n, qp, mode, layout, cols_first ? 1 : 0,
                  (const int*)dw.p, nullptr);
  else
    launch_fwd4x4((const int16_t*)di.p, (int16_t*)dout.p, n, qp, intra ? 1 : 0, mode, layout, nullptr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(out, dout.p, bytes, hipMemcpyDeviceToHost));
  return ICU_OK;
}

int icu_fwd4x4(const int16_t* in, int16_t* out, long n, int qp, int intra, int mode, int layout) {
  return tx_host(in, out, n, qp, intra, mode, layout, 0, 0, nullptr);
}

int icu_inv4x4(const int16_t* in, int16_t* out, long n, int qp, int mode, int layout, int cols_first,
               const int* w) {
  return tx_host(in, out, n, qp, 0, mode, layout, 1, cols_first, w);
}

// ---------------------------------------------------------------- batches
int icu_me_batch_create(const icu_me_config* cfgs, int n, icu_me_batch** out) {
  if (!cfgs || !out || n <= 0) return fail(ICU_E_ARG, "bad batch arguments");
  for (int i = 0; i < n; i++) {
    std::string why;
    if (!valid_cfg(cfgs[i], &why)) return fail(ICU_E_ARG, "stream " + std::to_string(i) + ": " + why);
    if (cfgs[i].predictor) return fail(ICU_E_ARG, "batched streams use the zero predictor");
    if (cfgs[i].metric != cfgs[0].metric) return fail(ICU_E_ARG, "batched streams share one metric");
  }
  icu_me_batch* b = new icu_me_batch();
  int rc = b->eng.init(cfgs, n);
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return ICU_OK;
}

int icu_me_batch_destroy(icu_me_batch* b) {
  delete b;
  return ICU_OK;
}

int icu_me_batch_num_mbs(icu_me_batch* b) { return b ? b->eng.total_mb : fail(ICU_E_ARG, "null batch"); }

int icu_me_batch_mb_base(icu_me_batch* b, int i) {
  if (!b || i < 0 || i >= (int)b->eng.sp.size()) return fail(ICU_E_ARG, "bad stream index");
  return b->eng.sp[i].mb_base;
}

int icu_me_batch_estimate_device(icu_me_batch* b, const uint8_t* const* curs, const uint8_t* const* refs,
                                 int16_t* d_mv, int32_t* d_dist, void* stream) {
  if (!b || !curs || !refs || !d_mv || !d_dist) return fail(ICU_E_ARG, "null argument");
  const hipStream_t st = pick(stream, b->eng);
  int slot = 0;
  int rc = b->eng.bind_eager(curs, refs, d_mv, d_dist, nullptr, st, &slot);
  if (rc) return rc;
  return b->eng.enqueue(slot, st);
}

int icu_me_batch_capture(icu_me_batch* b, const uint8_t* const* curs, const uint8_t* const* refs,
                         int16_t* d_mv, int32_t* d_dist) {
  if (!b || !curs || !refs || !d_mv || !d_dist) return fail(ICU_E_ARG, "null argument");
  Engine& e = b->eng;
  int rc = e.bind_graph(curs, refs, d_mv, d_dist);
  if (rc) return rc;
  if (e.gexec) hipGraphExecDestroy(e.gexec);
  if (e.graph) hipGraphDestroy(e.graph);
  e.gexec = nullptr;
  e.graph = nullptr;
  HIP_TRY(hipStreamBeginCapture(e.own, hipStreamCaptureModeThreadLocal));
  rc = e.enqueue(Engine::kGraph, e.own);
  hipGraph_t g = nullptr;
  hipError_t ce = hipStreamEndCapture(e.own, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (ce != hipSuccess) return fail(ICU_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
  e.graph = g;
  HIP_TRY(hipGraphInstantiate(&e.gexec, g, nullptr, nullptr, 0));
  return ICU_OK;
}

int icu_me_batch_graph_launch(icu_me_batch* b, void* stream) {
  if (!b || !b->eng.gexec) return fail(ICU_E_STATE, "no captured graph");
  HIP_TRY(hipGraphLaunch(b->eng.gexec, pick(stream, b->eng)));
  return ICU_OK;
}

int icu_me_batch_autotune(icu_me_batch* b, const uint8_t* const* curs, const uint8_t* const* refs, int16_t* d_mv,
                          int32_t* d_dist, int reps, float* best_ms) {
  if (!b || !curs || !refs || !d_mv || !d_dist) return fail(ICU_E_ARG, "null argument");
  Engine& e = b->eng;
  if (reps < 1) reps = 1;
  struct Plan {
    int mode, waves;
  };
  std::vector<Plan> plans;
  for (int m = 0; m < 3; m++) plans.push_back(Plan{m, e.queue_waves});
  if (e.any_wave_stream()) {
    plans.push_back(Plan{3, e.queue_waves});
    plans.push_back(Plan{5, e.queue_waves});
    for (int w : {6, 8}) {
      plans.push_back(Plan{4, w});
      plans.push_back(Plan{6, w});
    }
  }
  hipEvent_t ev[2] = {nullptr, nullptr};
  HIP_TRY(hipEventCreate(&ev[0]));
  if (hipEventCreate(&ev[1]) != hipSuccess) {
    hipEventDestroy(ev[0]);
    return fail(ICU_E_DEVICE, "hipEventCreate failed");
  }
  Plan best{e.grid_mode, e.queue_waves};
  float best_t = 1e30f;
  int rc = ICU_OK;
  // each plan is timed as the batch runs: captured in a graph and replayed
  auto run = [&]() { return icu_me_batch_graph_launch(b, e.own); };
  for (const Plan& p : plans) {
    e.grid_mode = p.mode;
    e.queue_waves = p.waves;
    if ((rc = e.replan()) != ICU_OK || (rc = icu_me_batch_capture(b, curs, refs, d_mv, d_dist)) != ICU_OK) break;
    for (int w = 0; w < 2 && rc == ICU_OK; w++) rc = run();  // warm-up
    // the least of three timed runs of `reps` estimates each
    float t_min = 1e30f;
    for (int trial = 0; trial < 3 && rc == ICU_OK; trial++) {
      if (hipEventRecord(ev[0], e.own) != hipSuccess) rc = fail(ICU_E_DEVICE, "hipEventRecord failed");
      for (int r = 0; r < reps && rc == ICU_OK; r++) rc = run();
      float ms = 0;
      if (rc == ICU_OK && (hipEventRecord(ev[1], e.own) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                           hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess))
        rc = fail(ICU_E_DEVICE, "event timing failed");
      t_min = std::min(t_min, ms / reps);
    }
    if (rc != ICU_OK) break;
    if (t_min < best_t) {
      best_t = t_min;
      best = p;
    }
  }
  hipEventDestroy(ev[0]);
  hipEventDestroy(ev[1]);
  // the graph of the last plan tried is dropped: the caller captures again
  if (hipStreamSynchronize(e.own) != hipSuccess && rc == ICU_OK) rc = fail(ICU_E_DEVICE, "tuning run failed");
  if (e.gexec) hipGraphExecDestroy(e.gexec);
  if (e.graph) hipGraphDestroy(e.graph);
  e.gexec = nullptr;
  e.graph = nullptr;
  if (rc != ICU_OK) return rc;
  e.grid_mode = best.mode;
  e.queue_waves = best.waves;
  if (best_ms) *best_ms = best_t;
  return e.replan();
}

int icu_me_batch_set_plan(icu_me_batch* b, int grid_mode, int queue_waves) {
  if (!b) return fail(ICU_E_ARG, "null batch");
  if (grid_mode < 0 || grid_mode > 6) return fail(ICU_E_ARG, "grid_mode must be 0..6");
  if (grid_mode >= 3 && !b->eng.any_wave_stream())
    return fail(ICU_E_ARG, "grid modes 3..6 need a stream the one-MB-per-wave kernels take");
  if (queue_waves < 1 || queue_waves > Engine::kMaxQueueWaves) return fail(ICU_E_ARG, "queue_waves must be 1..8");
  b->eng.grid_mode = grid_mode;
  b->eng.queue_waves = queue_waves;
  return b->eng.replan();
}

int icu_me_batch_get_plan(icu_me_batch* b, int* grid_mode, int* queue_waves) {
  if (!b) return fail(ICU_E_ARG, "null batch");
  if (grid_mode) *grid_mode = b->eng.grid_mode;
  if (queue_waves) *queue_waves = b->eng.queue_waves;
  return ICU_OK;
}

// ---------------------------------------------------------------- runtime
int icu_device_info(char* name, int name_len, int* cus, int* clock_khz) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, dev));
  if (name && name_len > 0) {
    std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (cus) *cus = p.multiProcessorCount;
  if (clock_khz) *clock_khz = p.clockRate;
  return ICU_OK;
}

void* icu_malloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    g_err = "hipMalloc failed";
    return nullptr;
  }
  return p;
}

int icu_free(void* d) {
  HIP_TRY(hipFree(d));
  return ICU_OK;
}

int icu_memcpy_htod(void* d, const void* h, size_t bytes, void* stream) {
  if (stream) {
    HIP_TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  } else {
    HIP_TRY(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
  }
  return ICU_OK;
}

int icu_memcpy_dtoh(void* h, const void* d, size_t bytes, void* stream) {
  if (stream) {
    HIP_TRY(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  } else {
    HIP_TRY(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
  }
  return ICU_OK;
}

int icu_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return ICU_OK;
}

int icu_memset(void* d, int value, size_t bytes, void* stream) {
  HIP_TRY(hipMemsetAsync(d, value, bytes, (hipStream_t)stream));
  return ICU_OK;
}

void* icu_stream_create(void) {
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
    g_err = "hipStreamCreate failed";
    return nullptr;
  }
  return (void*)s;
}

int icu_stream_destroy(void* s) {
  HIP_TRY(hipStreamDestroy((hipStream_t)s));
  return ICU_OK;
}

int icu_stream_sync(void* s) {
  HIP_TRY(hipStreamSynchronize((hipStream_t)s));
  return ICU_OK;
}

void* icu_event_create(void) {
  hipEvent_t e = nullptr;
  // timing events: a device-scope release when recorded, so recording one
  // does not force the caches' write-back to host scope between launches
  if (hipEventCreateWithFlags(&e, hipEventReleaseToDevice) != hipSuccess) {
    g_err = "hipEventCreate failed";
    return nullptr;
  }
  return (void*)e;
}

int icu_event_destroy(void* e) {
  HIP_TRY(hipEventDestroy((hipEvent_t)e));
  return ICU_OK;
}

int icu_event_record(void* e, void* s) {
  HIP_TRY(hipEventRecord((hipEvent_t)e, (hipStream_t)s));
  return ICU_OK;
}

int icu_event_elapsed_ms(void* a, void* b, float* ms) {
  HIP_TRY(hipEventSynchronize((hipEvent_t)b));
  HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
  return ICU_OK;
}

int icu_sqrt_check_device(const int* d_m, double* d_out, int n, void* stream) {
  if (!d_m || !d_out || n < 0) return fail(ICU_E_ARG, "bad arguments");
  launch_sqrt_check(d_m, d_out, n, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return ICU_OK;
}

}  // extern "C"
