// Host-side launchers shared between the translation units of the library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/icu_hip.h"
#include "icu_device.h"

struct icu_me;

namespace icu {

// Device bytes an estimator handle holds.
long long me_device_bytes(const ::icu_me* h);
// Device address of the handle's median-predictor wavefront error flag (one
// int, nonzero after a bounded wait ran out), or nullptr without a predictor.
int* me_wave_flag(::icu_me* h);
// Reads and clears that flag (the work that sets it must be done):
// ICU_E_STATE if it was set.
int me_check(::icu_me* h);

// Records msg as the calling thread's last error and returns code.
int icu_fail(int code, const std::string& msg);

// Search tiling of one stream: V candidate rows per lane tile, B row tiles,
// Q column quads, G MBs per workgroup task.
struct Tiling {
  int V, B, Q, G;
};
// The planner's tiling for a picture mbw MBs wide and window half-width L;
// argmin: no cost table (exact lane argmin keys needed).  V_req / G_req
// nonzero force those; V = 0 when a forced V is not allowed.
Tiling plan_tiling(int mbw, int L, int metric, bool argmin, int V_req, int G_req);
// Configuration checks of icu_me_create; *why receives the reason.
bool valid_cfg(const icu_me_config& c, std::string* why);

// Per-stream parameters of the decision / sub-pel stage.
struct DecideDesc {
  const uint8_t* cur;
  const uint8_t* ref;  // padded plane, pointer at pixel (0,0)
  int cur_stride, ref_stride;
  int W, H, mbw, mbh;
  int R, L;
  int metric, filter, subpel, pred, bestvec;
  int mb_base;                 // first MB slot of this stream in ires / outputs
  const uint32_t* table;       // cost table (nullptr: use ires)
  long long table_base;
  int table_dim;
  const IntResult* ires;
  int16_t* mv_out;             // 2 per MB slot
  int32_t* dist_out;           // 1 per MB slot
  int32_t* int_out;            // 3 per MB slot (integer stage), may be null
  int* ticket;                 // [0] row ticket, [1] error flag
};

// Work item of the parallel decision kernel.
struct DecideItem {
  int stream;
  int mb;
};

// One reference plane to pad: an 8-bit W x H plane (stride src_stride) into
// the padded plane whose pixel (0,0) is dst (stride dst_stride).  wg_base:
// the plane's first workgroup in the shared launch.
struct PadDesc {
  const uint8_t* src;
  uint8_t* dst;
  int W, H, src_stride, dst_stride;
  int wg_base;
};
// Workgroups padding one W x H plane.
int pad_plane_workgroups(int W, int H);
// Pads n planes in one launch of total_wgs workgroups.
void launch_pad_planes(const PadDesc* d_pads, int n, int total_wgs, hipStream_t st);
void launch_search(const StreamDesc* d_streams, const SearchTask* d_tasks, int ntasks, int grid_mode,
                   int metric, int vfix, bool table, int buf_dwords, IntResult* d_out,
                   uint32_t* d_table, const DecideDesc* d_decide, bool band, hipStream_t st);
// One MB per wave (me_wave_kernel) for argmin SAD streams whose window
// half-width L is L0 or L0 - 1, L0 = wave_class(L) in {8, 16, 32, 64} (0: no
// class); decides each MB in place.  All tasks of one launch share L0.
// lean: every stream of the launch is of the lean modes (wave_lean), and
// the kernel carries only their decision.
int wave_class(int L);
void launch_wave_search(const StreamDesc* d_streams, const SearchTask* d_tasks, int ntasks, int L0,
                        const DecideDesc* d_decide, bool band, bool lean, hipStream_t st);
// A wave-class stream whose decision is integer or half pel with no final
// vector choice (GetBestVector off).
inline bool wave_lean(int subpel, int best_vector) { return subpel <= 1 && !best_vector; }
// The same MBs, any mix of classes, in one persistent launch of `waves`
// one-wave workgroups: wave b runs tasks [d_woff[b], d_woff[b + 1]) in order
// (d_woff: waves + 1 offsets).  big_only: every task is of class 32 or 64
// (the kernel then carries only that search path).
void launch_wave_queue(const StreamDesc* d_streams, const SearchTask* d_tasks, const int* d_woff, int waves,
                       bool big_only, bool lean, const DecideDesc* d_decide, hipStream_t st);
// Estimated work of the MB at pel (n, m) of a W x H picture searched by the
// wave kernels with window half-width L, in candidates (the queue's LPT key).
long wave_task_cost(int n, int m, int W, int H, int L);
// Dwords of one staging buffer (window rows padded to 16 bytes + blocks).
int search_stage_dwords(int B, int V, int Q, int G);
int search_max_window_dwords();
int search_max_group();

void launch_decide_parallel(const DecideDesc* d_descs, const DecideItem* d_items, int nitems,
                            hipStream_t st);
// Median-predictor streams: the predictor-free pass of the stream's nmb MBs
// into d_pf (predfree_bytes() each), then the wavefront that finishes each
// MB once its neighbours are decided.  table_dim: the per-MB cost table side
// (0 when no table is kept).
int predfree_bytes();
void launch_decide_predfree(const DecideDesc* d_desc, void* d_pf, int nmb, hipStream_t st);
void launch_decide_wavefront(const DecideDesc* d_desc, int table_dim, int mbh, const void* d_pf,
                             hipStream_t st);
void launch_fill_i32(int* p, int v, int n, hipStream_t st);
// Adds an order-free checksum of `bytes` bytes at p (4-byte aligned) to *out.
void launch_checksum(const void* p, long bytes, unsigned long long* out, hipStream_t st);

// Multi-reference variable-block-size search (me_vbs.hip).
constexpr int kVbsMaxRefs = 4;
struct VbsParams {
  const uint8_t* cur;                  // W x H, stride W
  const uint8_t* ref[kVbsMaxRefs];     // padded planes at pixel (0,0)
  int ref_stride;
  int W, H, mbw, nmb;
  int R, N, nrefs, lambda;
  int dx0, Q, shift, V, Gr, rows, row_dw;  // tile plan (vbs_plan)
  int wave;                            // 8 or 16: one MB per wave (vbs_wave_kernel<wave>); 0: workgroups
  uint8_t* mode;                       // per MB
  int16_t* mv8;                        // 2 per 8x8 block, quarter pel
  uint8_t* ref8;                       // per 8x8 block
  int32_t* cost;                       // per MB
};
// Fills the tile plan fields of p for range R and nrefs references.
void vbs_plan(int R, int nrefs, VbsParams* p);
// The one-MB-per-wave class of range R: 8 or 16 for R = 7, 8, 15, 16, else 0.
int vbs_wave_class(int R);
int vbs_lds_bytes(const VbsParams& p);
int vbs_tile_rows(const VbsParams& p);
void launch_vbs_search(const VbsParams& p, hipStream_t st);
// Pads up to kVbsMaxRefs planes in one launch, descriptors passed by value.
struct PadSet {
  PadDesc d[kVbsMaxRefs];
  int n;
};
void launch_pad_set(const PadSet& s, int total_wgs, hipStream_t st);

// Transform / coding launchers (it_kernels.hip).
void launch_fwd4x4(const int16_t* in, int16_t* out, long n, int qp, int intra, int mode,
                   int layout, hipStream_t st);
void launch_inv4x4(const int16_t* in, int16_t* out, long n, int qp, int mode, int layout,
                   int cols_first, const int* d_weights, hipStream_t st);
// QPs a frame's residual is coded at in one pass, with an output pair each.
constexpr int kMaxQpSet = 20;
struct QpSet {
  int n;
  int qp[kMaxQpSet];
  int16_t* levels[kMaxQpSet];
  uint8_t* recon[kMaxQpSet];
};
void launch_code_frame(const uint8_t* cur, int cur_stride, const uint8_t* ref_padded,
                       int ref_stride, int W, int H, const int16_t* mv, const QpSet& qs, int intra,
                       int cols_first, int16_t* pred_out, hipStream_t st);
// Residual coding of a frame predicted per 8x8 block from one of up to
// kVbsMaxRefs padded references at an integer vector (mv8 in quarter pel,
// multiples of 4; ref8 the reference index).
struct VbsRefs {
  const uint8_t* ref[kVbsMaxRefs];
  int stride;
};
void launch_code_vbs(const uint8_t* cur, const VbsRefs& refs, int W, int H, const int16_t* mv8,
                     const uint8_t* ref8, const QpSet& qs, int intra, int cols_first,
                     int16_t* pred_out, hipStream_t st);
// Whole-frame YUV 4:2:0 compensation from padded planes (luma stride ys,
// chroma stride cs, pointers at pixel (0,0)) into 8-bit output planes.
void launch_mc_yuv(const uint8_t* ref_y, int ys, const uint8_t* ref_u, const uint8_t* ref_v, int cs, int W,
                   int H, const int16_t* mv, uint8_t* out_y, uint8_t* out_u, uint8_t* out_v, hipStream_t st);
void launch_sqrt_check(const int* m, double* out, int n, hipStream_t st);

}  // namespace icu
