// Serving loop: frame pairs from many streams, admitted as they arrive,
// estimated and coded on the device with their work overlapping.
//
// Continuous batching over HIP streams ("lanes"): a submitted frame is
// queued per stream and admitted as soon as (a) its stream has no frame in
// flight (the frame's reference is the previous frame's reconstruction, kept
// resident in HBM), (b) fewer than max_inflight frames are on the device and
// (c) its transient buffers fit the device-memory budget.  Admitted frames
// are spread over the lanes, so frames of different streams run at once; a
// frame that arrives while others run starts when a slot frees up, not when
// a batch completes.  Each frame: upload, motion estimation
// (icu_me_estimate_device), luma MC + residual coding at the server's QP
// (icu_code_frame_device), checksums of the levels and reconstruction, and a
// download of the vectors.  No reference counterpart: the reference codes one
// frame per call on the caller's thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/icu_hip.h"
#include "icu_internal.h"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Free lists of same-sized buffers, so steady-state admission allocates
// nothing.  `pinned`: page-locked host memory (asynchronous copies).
struct Pool {
  bool pinned = false;
  std::map<size_t, std::vector<void*>> free;
  std::vector<void*> all;
  // `pref`: a buffer to hand out again if it is free (keeps a stream's
  // estimator bound to the same buffers from frame to frame).
  void* get(size_t bytes, void* pref = nullptr) {
    auto& f = free[bytes];
    if (pref) {
      auto it = std::find(f.begin(), f.end(), pref);
      if (it != f.end()) {
        f.erase(it);
        return pref;
      }
    }
    if (!f.empty()) {
      void* p = f.back();
      f.pop_back();
      return p;
    }
    void* p = nullptr;
    hipError_t e = pinned ? hipHostMalloc(&p, bytes, hipHostMallocDefault) : hipMalloc(&p, bytes);
    if (e != hipSuccess) return nullptr;
    all.push_back(p);
    return p;
  }
  void put(void* p, size_t bytes) { free[bytes].push_back(p); }
  void release() {
    for (void* p : all) pinned ? (void)hipHostFree(p) : (void)hipFree(p);
    all.clear();
    free.clear();
  }
};

// Frames of a stream whose transient buffers are allocated when it opens.
constexpr int kPoolFrames = 3;

struct Pending {
  long long tag;
  void* h_frame;  // pinned copy of the submitted frame
  double submit_s;
  int frame;
};

struct SrvStream {
  icu_me* h = nullptr;
  int W = 0, H = 0, nmb = 0;
  uint8_t* d_ref = nullptr;    // previous frame's reconstruction
  uint8_t* d_recon = nullptr;  // this frame's reconstruction
  int16_t* d_levels = nullptr;
  unsigned long long* d_sums = nullptr;  // [0] levels, [1] reconstruction
  long long resident = 0;
  void* last[3] = {};  // transient buffers of the previous frame
  int submitted = 0, done = 0;
  bool busy = false;
  std::deque<Pending> queue;
};

struct Flight {
  int stream;
  Pending p;
  hipEvent_t ev;
  uint8_t* d_cur;
  int16_t* d_mv;
  int32_t* d_dist;
  void* h_out;  // pinned: vectors (4 B per MB), the two checksums, the wavefront error flag
  long long transient;
};

}  // namespace

struct icu_server {
  icu_server_config cfg{};
  std::vector<hipStream_t> lanes;
  int next_lane = 0;
  int next_admit = 0;
  std::vector<SrvStream> streams;
  std::vector<Flight> inflight;
  std::vector<hipEvent_t> events;
  Pool dev, host;
  long long bytes = 0, peak = 0;  // resident + in flight
  long long resident = 0;         // streams' resident state
  int queued = 0;
  std::vector<std::vector<int16_t>> out_mv;  // result vectors, valid until the next poll

  ~icu_server() {
    for (hipStream_t s : lanes) hipStreamSynchronize(s);
    for (auto& f : inflight) hipEventDestroy(f.ev);
    for (hipEvent_t e : events) hipEventDestroy(e);
    for (auto& s : streams) {
      if (s.h) icu_me_destroy(s.h);
      hipFree(s.d_ref);
      hipFree(s.d_recon);
      hipFree(s.d_levels);
      hipFree(s.d_sums);
    }
    dev.release();
    host.release();
    for (hipStream_t s : lanes) hipStreamDestroy(s);
  }

  static size_t out_bytes(const SrvStream& s) { return 4 * (size_t)s.nmb + 24; }

  static long long transient_bytes(const SrvStream& s) {
    return (long long)s.W * s.H + 8LL * s.nmb;  // current plane, vectors, distortions
  }

  int launch(int si) {
    SrvStream& s = streams[si];
    Pending p = s.queue.front();
    Flight f{};
    f.stream = si;
    f.p = p;
    f.transient = transient_bytes(s);
    f.d_cur = (uint8_t*)dev.get((size_t)s.W * s.H, s.last[0]);
    f.d_mv = (int16_t*)dev.get(4 * (size_t)s.nmb, s.last[1]);
    f.d_dist = (int32_t*)dev.get(4 * (size_t)s.nmb, s.last[2]);
    s.last[0] = f.d_cur;
    s.last[1] = f.d_mv;
    s.last[2] = f.d_dist;
    f.h_out = host.get(out_bytes(s));
    if (!f.d_cur || !f.d_mv || !f.d_dist || !f.h_out) return icu::icu_fail(ICU_E_DEVICE, "server buffer allocation failed");
    if (events.empty()) {
      hipEvent_t e;
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
        return icu::icu_fail(ICU_E_DEVICE, "hipEventCreate failed");
      events.push_back(e);
    }
    f.ev = events.back();
    events.pop_back();
    hipStream_t st = lanes[next_lane];
    next_lane = (next_lane + 1) % (int)lanes.size();
    const size_t px = (size_t)s.W * s.H;
    if (hipMemcpyAsync(f.d_cur, p.h_frame, px, hipMemcpyHostToDevice, st) != hipSuccess)
      return icu::icu_fail(ICU_E_DEVICE, "frame upload failed");
    // the first frame of a stream is its own reference
    const uint8_t* ref = p.frame == 0 ? f.d_cur : s.d_ref;
    int rc = icu_me_estimate_device(s.h, f.d_cur, ref, f.d_mv, f.d_dist, nullptr, st);
    if (!rc)
      rc = icu_code_frame_device(s.h, f.d_cur, f.d_mv, cfg.qp, 0, 0, s.d_levels, s.d_recon, nullptr, st);
    if (rc) return rc;
    icu::launch_fill_i32((int*)s.d_sums, 0, 4, st);
    icu::launch_checksum(s.d_levels, 2 * (long)px, s.d_sums, st);
    icu::launch_checksum(s.d_recon, (long)px, s.d_sums + 1, st);
    // median-predictor streams: the wavefront's error flag comes back with
    // the results (it stays set on the device until read)
    int* flag = icu::me_wave_flag(s.h);
    int* h_flag = (int*)((char*)f.h_out + 4 * (size_t)s.nmb + 16);
    *h_flag = 0;
    if (hipMemcpyAsync(f.h_out, f.d_mv, 4 * (size_t)s.nmb, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync((char*)f.h_out + 4 * (size_t)s.nmb, s.d_sums, 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
        (flag && hipMemcpyAsync(h_flag, flag, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipEventRecord(f.ev, st) != hipSuccess)
      return icu::icu_fail(ICU_E_DEVICE, "result download failed");
    s.queue.pop_front();
    queued--;
    s.busy = true;
    bytes += f.transient;
    peak = std::max(peak, bytes);
    inflight.push_back(f);
    return ICU_OK;
  }

  // Admit queued frames, streams in rotating order, while slots and the
  // budget allow.
  int admit() {
    const int n = (int)streams.size();
    for (int k = 0; k < n && (int)inflight.size() < cfg.max_inflight; k++) {
      const int si = (next_admit + k) % n;
      SrvStream& s = streams[si];
      if (s.busy || s.queue.empty()) continue;
      if (bytes + transient_bytes(s) > cfg.mem_budget) continue;
      int rc = launch(si);
      if (rc) return rc;
      next_admit = (si + 1) % n;
    }
    return ICU_OK;
  }

  int collect(icu_server_result* out, int max_out, int* n) {
    for (size_t i = 0; i < inflight.size() && *n < max_out;) {
      Flight& f = inflight[i];
      const hipError_t q = hipEventQuery(f.ev);
      if (q == hipErrorNotReady) {
        i++;
        continue;
      }
      if (q != hipSuccess) return icu::icu_fail(ICU_E_DEVICE, std::string("frame failed: ") + hipGetErrorString(q));
      SrvStream& s = streams[f.stream];
      icu_server_result& r = out[(*n)++];
      out_mv.emplace_back((const int16_t*)f.h_out, (const int16_t*)f.h_out + 2 * (size_t)s.nmb);
      const unsigned long long* sums = (const unsigned long long*)((const char*)f.h_out + 4 * (size_t)s.nmb);
      const int wave_error = *(const int*)((const char*)f.h_out + 4 * (size_t)s.nmb + 16);
      r.tag = f.p.tag;
      r.stream = f.stream;
      r.frame = f.p.frame;
      r.nmb = s.nmb;
      r.submit_s = f.p.submit_s;
      r.done_s = now_s();
      r.levels_sum = sums[0];
      r.recon_sum = sums[1];
      r.mv = out_mv.back().data();
      std::swap(s.d_ref, s.d_recon);  // this reconstruction is the next frame's reference
      s.busy = false;
      s.done++;
      dev.put(f.d_cur, (size_t)s.W * s.H);
      dev.put(f.d_mv, 4 * (size_t)s.nmb);
      dev.put(f.d_dist, 4 * (size_t)s.nmb);
      host.put(f.h_out, out_bytes(s));
      host.put(f.p.h_frame, (size_t)s.W * s.H);
      bytes -= f.transient;
      events.push_back(f.ev);
      const int stream = f.stream, frame = f.p.frame;
      inflight.erase(inflight.begin() + i);
      if (wave_error) {
        // the frame's vectors used neighbours a timed-out wait never saw:
        // report it (me_check clears the device flag)
        (*n)--;
        out_mv.pop_back();
        icu::me_check(s.h);
        return icu::icu_fail(ICU_E_STATE, "median-predictor wavefront timed out (stream " +
                                              std::to_string(stream) + ", frame " + std::to_string(frame) + ")");
      }
    }
    return ICU_OK;
  }
};

extern "C" {

double icu_server_clock(void) { return now_s(); }

int icu_server_create(const icu_server_config* cfg, icu_server** out) {
  if (!cfg || !out) return icu::icu_fail(ICU_E_ARG, "null argument");
  if (cfg->lanes < 1 || cfg->lanes > 16) return icu::icu_fail(ICU_E_ARG, "lanes must be 1..16");
  if (cfg->max_inflight < 1) return icu::icu_fail(ICU_E_ARG, "max_inflight must be >= 1");
  if (cfg->mem_budget <= 0) return icu::icu_fail(ICU_E_ARG, "mem_budget must be positive");
  if (cfg->qp < 0 || cfg->qp > 51) return icu::icu_fail(ICU_E_ARG, "qp must be 0..51");
  icu_server* s = new icu_server();
  s->cfg = *cfg;
  s->host.pinned = true;
  for (int i = 0; i < cfg->lanes; i++) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      delete s;
      return icu::icu_fail(ICU_E_DEVICE, "hipStreamCreate failed");
    }
    s->lanes.push_back(st);
  }
  *out = s;
  return ICU_OK;
}

int icu_server_destroy(icu_server* s) {
  delete s;
  return ICU_OK;
}

int icu_server_open_stream(icu_server* s, const icu_me_config* cfg, int* stream_id) {
  if (!s || !cfg || !stream_id) return icu::icu_fail(ICU_E_ARG, "null argument");
  const long long px = (long long)cfg->width * cfg->height;
  SrvStream st;
  st.W = cfg->width;
  st.H = cfg->height;
  st.nmb = (int)(px / 256);
  int rc = icu_me_create(cfg, &st.h);
  if (rc) return rc;
  // resident: the estimator handle, reference + reconstruction, levels, checksums
  const long long resident = icu::me_device_bytes(st.h) + 2 * px + 2 * px + 32;
  if (s->bytes + resident > s->cfg.mem_budget) {
    icu_me_destroy(st.h);
    return icu::icu_fail(ICU_E_RANGE, "stream does not fit the server's device-memory budget");
  }
  if (hipMalloc(&st.d_ref, px) != hipSuccess || hipMalloc(&st.d_recon, px) != hipSuccess ||
      hipMalloc(&st.d_levels, 2 * px) != hipSuccess || hipMalloc(&st.d_sums, 32) != hipSuccess) {
    icu_me_destroy(st.h);
    hipFree(st.d_ref);
    hipFree(st.d_recon);
    hipFree(st.d_levels);
    hipFree(st.d_sums);
    return icu::icu_fail(ICU_E_DEVICE, "stream allocation failed");
  }
  // fill the buffer pools for a few frames of this stream now: growing a
  // pool mid-run calls hipMalloc / hipHostMalloc, which block the caller for
  // milliseconds
  {
    const size_t ob = icu_server::out_bytes(st);
    void* d[3 * kPoolFrames];
    void* h[2 * kPoolFrames];
    bool ok = true;
    for (int k = 0; k < kPoolFrames; k++) {
      d[3 * k] = s->dev.get((size_t)px);
      d[3 * k + 1] = s->dev.get(4 * (size_t)st.nmb);
      d[3 * k + 2] = s->dev.get(4 * (size_t)st.nmb);
      h[2 * k] = s->host.get((size_t)px);
      h[2 * k + 1] = s->host.get(ob);
      ok = ok && d[3 * k] && d[3 * k + 1] && d[3 * k + 2] && h[2 * k] && h[2 * k + 1];
    }
    for (int k = 0; k < kPoolFrames; k++) {
      if (d[3 * k]) s->dev.put(d[3 * k], (size_t)px);
      if (d[3 * k + 1]) s->dev.put(d[3 * k + 1], 4 * (size_t)st.nmb);
      if (d[3 * k + 2]) s->dev.put(d[3 * k + 2], 4 * (size_t)st.nmb);
      if (h[2 * k]) s->host.put(h[2 * k], (size_t)px);
      if (h[2 * k + 1]) s->host.put(h[2 * k + 1], ob);
    }
    if (!ok) {
      icu_me_destroy(st.h);
      hipFree(st.d_ref);
      hipFree(st.d_recon);
      hipFree(st.d_levels);
      hipFree(st.d_sums);
      return icu::icu_fail(ICU_E_DEVICE, "stream buffer pool allocation failed");
    }
  }
  st.resident = resident;
  s->resident += resident;
  s->bytes += resident;
  s->peak = std::max(s->peak, s->bytes);
  s->streams.push_back(st);
  *stream_id = (int)s->streams.size() - 1;
  return ICU_OK;
}

int icu_server_submit(icu_server* s, int stream_id, const uint8_t* frame, long long tag) {
  if (!s || !frame) return icu::icu_fail(ICU_E_ARG, "null argument");
  if (stream_id < 0 || stream_id >= (int)s->streams.size()) return icu::icu_fail(ICU_E_ARG, "bad stream id");
  SrvStream& st = s->streams[stream_id];
  if (s->resident + icu_server::transient_bytes(st) > s->cfg.mem_budget)
    return icu::icu_fail(ICU_E_RANGE, "a frame of this stream can never fit the device-memory budget");
  const size_t px = (size_t)st.W * st.H;
  void* h = s->host.get(px);
  if (!h) return icu::icu_fail(ICU_E_DEVICE, "pinned staging allocation failed");
  std::memcpy(h, frame, px);
  st.queue.push_back(Pending{tag, h, now_s(), st.submitted++});
  s->queued++;
  return s->admit();
}

int icu_server_poll(icu_server* s, icu_server_result* out, int max_out, int wait) {
  if (!s || (!out && max_out > 0) || max_out < 0) return icu::icu_fail(ICU_E_ARG, "bad arguments");
  s->out_mv.clear();
  s->out_mv.reserve(max_out);
  int n = 0;
  int rc = s->collect(out, max_out, &n);
  if (!rc) rc = s->admit();
  if (!rc && wait && n == 0 && !s->inflight.empty() && max_out > 0) {
    if (hipEventSynchronize(s->inflight.front().ev) != hipSuccess)
      return icu::icu_fail(ICU_E_DEVICE, "wait for a frame failed");
    rc = s->collect(out, max_out, &n);
    if (!rc) rc = s->admit();
  }
  return rc ? rc : n;
}

int icu_server_stream_estimator(icu_server* s, int stream_id, icu_me** out) {
  if (!s || !out) return icu::icu_fail(ICU_E_ARG, "null argument");
  if (stream_id < 0 || stream_id >= (int)s->streams.size()) return icu::icu_fail(ICU_E_ARG, "bad stream id");
  *out = s->streams[stream_id].h;
  return ICU_OK;
}

int icu_server_stats(icu_server* s, long long* bytes_in_use, long long* peak_bytes, int* queued,
                     int* inflight) {
  if (!s) return icu::icu_fail(ICU_E_ARG, "null server");
  if (bytes_in_use) *bytes_in_use = s->bytes;
  if (peak_bytes) *peak_bytes = s->peak;
  if (queued) *queued = s->queued;
  if (inflight) *inflight = (int)s->inflight.size();
  return ICU_OK;
}

}  // extern "C"
