/*
 * icu_hip.h - C ABI of the MI355X (gfx950) motion-estimation and H.264 4x4
 * transform/quant library.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns 0 on success or a negative ICU_E* code; icu_last_error() then holds
 * a message for the calling thread.  Host-buffer entry points are
 * synchronous.  *_device entry points take device pointers and a HIP stream
 * (hipStream_t passed as void*, NULL = the handle's own stream) and are
 * asynchronous on that stream.
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef ICU_HIP_H
#define ICU_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ICU_OK 0
#define ICU_E_ARG -1      /* invalid argument or configuration */
#define ICU_E_RANGE -2    /* input sample outside 0..255 */
#define ICU_E_DEVICE -3   /* HIP runtime / allocation failure */
#define ICU_E_STATE -4    /* handle not ready, or a kernel reported failure */

/* Metric, window and mode switches of the estimator.  The first value of
 * each is what the reference class ships with. */
#define ICU_METRIC_SSD 0 /* MotionEstimatorH264ImplFull.cpp:74 (#undef MEH264IF_ABS_DIFF) */
#define ICU_METRIC_SAD 1 /* the MEH264IF_ABS_DIFF build */
#define ICU_WINDOW_REFERENCE 0 /* +-(R-1): GetMotionRange, Full.cpp:399 -> :1029-1081 */
#define ICU_WINDOW_FULL 1      /* +-R, same 16-pel picture-border clamp */

typedef struct icu_me_config {
  int width, height; /* luma size, multiples of 16 */
  int range;         /* full-pel search range R (the reference's motionRange / 4), 1..64 */
  int metric;        /* ICU_METRIC_SSD | ICU_METRIC_SAD */
  int window;        /* ICU_WINDOW_REFERENCE | ICU_WINDOW_FULL */
  int filter;        /* 1: reference running-best candidate filter (Full.cpp:410-411); 0: off */
  int subpel;        /* 0 integer, 1 half, 2 half then quarter (reference) */
  int predictor;     /* 0 zero vector, 1 H.264 median of A/B/C(D) (H264MotionVectorPredictorImpl1) */
  int best_vector;   /* 1: GetBestVector over {best, predictor, zero} (reference); 0: keep best */
  int tile_rows;     /* 0: chosen by the tuner; else candidate rows per lane: 8, 11, 13 or 16 */
  int mbs_per_task;  /* 0: chosen by the tuner; else MBs staged per workgroup task (1..16) */
  int grid_mode;     /* 0: the planner's choice (one MB per wave where the window takes it, else one
                        task per workgroup) until tuned; else 1 + the mode icu_me_get_tiling
                        reports (1 one task per workgroup, 2 persistent, 3 persistent balanced,
                        4 one MB per wave, a launch per window class, 5 one MB per wave in one
                        persistent launch over per-XCD queues, 6 as 4 with the +-8 / +-16
                        classes on a second stream, beside the others, 7 the +-31/32 and +-63/64
                        classes as in 5 and the others as in 6); a batch takes stream 0's */
} icu_me_config;

typedef struct icu_me icu_me;

/* Fill cfg with the reference class's behaviour for a W x H picture and a
 * quarter-pel motionRange, as MotionEstimatorH264ImplFull(pSrc, pRef, W, H,
 * motionRange, pMVPred) with an H264MotionVectorPredictorImpl1 predictor
 * (include/CodecUtils/MotionEstimatorH264ImplFull.h:77-90). */
int icu_me_config_reference(icu_me_config* cfg, int width, int height, int motion_range_qpel);

/* Create / destroy an estimator: replaces the constructor + Create() and the
 * destructor of MotionEstimatorH264ImplFull
 * (src/CodecUtils/MotionEstimatorH264ImplFull.cpp:111-153,216-320). */
int icu_me_create(const icu_me_config* cfg, icu_me** out);
int icu_me_destroy(icu_me* h);

/* Estimate one frame pair from host int16 planes (W*H samples each, 0..255):
 * replaces IMotionEstimator::Estimate(long* avgDistortion)
 * (include/CodecUtils/IMotionEstimator.h:86-96, Full.cpp:330-611).
 * mv_out: 2 int16 per MB (x, y) in quarter pels, raster order - the
 * VectorStructList SIMPLE2D layout (src/CodecUtils/VectorStructList.cpp:120-196).
 * dist_out (nullable): distortion of each chosen vector.  include (nullable):
 * one flag per MB; avg_out receives sum(dist of included) / count, or 0
 * without a map, as the reference's avgDistortion. */
int icu_me_estimate(icu_me* h, const int16_t* cur, const int16_t* ref, const uint8_t* include,
                    int16_t* mv_out, int32_t* dist_out, long* avg_out);

/* Same on device-resident 8-bit planes (stride = width).  d_mv / d_dist as
 * above, device memory; d_int (nullable) receives the integer-stage
 * (mx, my, cost) per MB. */
int icu_me_estimate_device(icu_me* h, const uint8_t* d_cur, const uint8_t* d_ref, int16_t* d_mv,
                           int32_t* d_dist, int32_t* d_int, void* stream);

/* icu_me_estimate_device without a handle and without allocating: no
 * hipMalloc / hipFree or other allocating call on any path.  No counterpart
 * in the reference; it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, hosts that forbid allocation).  The
 * padded reference, descriptors and task list live in a static device
 * scratch array of 20 MiB per GPU, held until the process ends; a picture
 * whose padded plane and tasks exceed it (beyond about 3840x2160) gets
 * ICU_E_RANGE.  cfg as icu_me_create, with filter 0 and predictor 0 (the
 * modes that keep no per-MB cost table; ICU_E_ARG otherwise).  Outputs as
 * icu_me_estimate_device, asynchronous on stream (NULL: the null stream).
 * All calls on one GPU share its scratch: calls that overlap on one GPU
 * (other streams, other threads) race. */
int icu_me_estimate_device_noalloc(const icu_me_config* cfg, const uint8_t* d_cur, const uint8_t* d_ref,
                                   int16_t* d_mv, int32_t* d_dist, int32_t* d_int, void* stream);

/* icu_me_estimate_device with HIP events between its stages, recorded on
 * the stream the kernels run on; waits for completion and writes the
 * duration of each stage in ms: [0] reference padding, [1] integer search
 * (with the decision when no cost table is kept), [2] decision kernels (0
 * when the search kernel decides every MB itself).
 * search_reps (1..1000) back-to-back search launches share one event pair
 * and [1] is their mean, which takes the events' own dispatch gap out of a
 * single launch's time (the search rewrites the same results each time).
 * No reference counterpart (the reference has no timers); for benchmarks. */
int icu_me_estimate_timed(icu_me* h, const uint8_t* d_cur, const uint8_t* d_ref, int16_t* d_mv,
                          int32_t* d_dist, void* stream, int search_reps, float* stage_ms);

/* Waits for the device and reports whether a median-predictor wavefront of
 * this handle ran out its bounded wait for a neighbour row since the last
 * check (ICU_E_STATE; its vectors are then not the reference's).  The flag
 * stays set on the device until read here (icu_me_estimate reads it too);
 * while it is set, later wavefronts of the handle stop waiting as well.
 * Asynchronous callers of icu_me_estimate_device with the median predictor
 * call this after their results are complete.  No reference counterpart
 * (the reference's raster loop cannot time out). */
int icu_me_check(icu_me* h);
/* Median-predictor handles: waits for the device and counts the MBs of the
 * last estimate whose integer scan met an equal cost, so the wavefront
 * decided them again with their predictor (the others only add the
 * predictor's distortion and the final choice there).  0 for handles
 * without the predictor.  No reference counterpart; a diagnostic. */
int icu_me_tie_count(icu_me* h, int* tie_mbs);
/* Test hook: sets the handle's wavefront error flag from the host, so the
 * reporting paths can be exercised without stalling a GPU.  ICU_E_STATE for
 * a handle without the median predictor.  No reference counterpart. */
int icu_me_debug_set_wave_error(icu_me* h);

/* Tile choice in effect for this handle (after tuning): candidate rows per
 * lane, MBs per workgroup task, and the search grid mode: 0 one task per
 * workgroup, 1 persistent (as many workgroups as stay resident claim tasks
 * until none are left), 2 persistent with the grid reduced so every
 * workgroup runs equally many tasks, 3 one MB per wave (SAD without a cost
 * table at +-31 / +-32: the block in scalar registers, 64 x 64 candidates
 * in lane tiles plus the last column and row; the rows per lane and MBs
 * per task reported then apply to no stream of a single-stream handle). */
int icu_me_get_tiling(icu_me* h, int* tile_rows, int* mbs_per_task, int* grid_mode);

/* Time each candidate tiling (rows per lane x MBs per task x grid mode) on
 * the given device frames and keep the fastest (the tuner).  reps: timed
 * repetitions per candidate.  No reference counterpart. */
int icu_me_autotune(icu_me* h, const uint8_t* d_cur, const uint8_t* d_ref, int reps);

/* Luma motion compensation + residual coding of one frame on the device:
 * prediction from the handle's padded reference (the last estimated frame)
 * at the vectors d_mv (MotionCompensatorH264ImplStd::Compensate,
 * src/CodecUtils/MotionCompensatorH264ImplStd.cpp:323-437), residual
 * (OverlayMem2Dv2::Sub16x16), forward IT+Q (FastForward4x4ITImpl1), levels
 * out, dequant + inverse IT (rows first: FastInverse4x4ITImpl1; cols first:
 * FastInverse4x4On16x16ITImpl1) and clipped reconstruction
 * (OverlayMem2Dv2::Add16x16WithClip255).  Levels and pred are int16 W*H in
 * picture layout, recon 8-bit W*H.  d_pred is nullable. */
int icu_code_frame_device(icu_me* h, const uint8_t* d_cur, const int16_t* d_mv, int qp, int intra,
                          int cols_first, int16_t* d_levels, uint8_t* d_recon, int16_t* d_pred,
                          void* stream);

/* icu_code_frame_device at nqp (1..20) QPs in one pass: the prediction and
 * residual are formed once, then coded and reconstructed at each qps[i] into
 * d_levels[i] / d_recon[i] (host arrays of device pointers).  No reference
 * counterpart: the reference codes a frame at one QP per call; this is the
 * same per-QP computation, batched. */
int icu_code_frame_qps_device(icu_me* h, const uint8_t* d_cur, const int16_t* d_mv, int nqp,
                              const int* qps, int intra, int cols_first, int16_t* const* d_levels,
                              uint8_t* const* d_recon, int16_t* d_pred, void* stream);

/* Batched 4x4 transforms on host buffers.  layout 0: contiguous 4x4 blocks
 * (16 int16 each, raster within the block); layout 1: 16x16 raster MBs of
 * 16 blocks (the FastForward4x4On16x16ITImpl1 layout).  mode 0 IT+Q, 1 IT
 * only, 2 Q only (IForwardTransform::SetMode).  Up to 16384 blocks are
 * transformed on the CPU with the device kernels' arithmetic (a per-block
 * codec call must not pay a device round trip); larger batches run on the
 * device.  in and out may be the same buffer.
 * icu_fwd4x4 replaces IForwardTransform::Transform + SetParameter(QUANT_ID,
 *   INTRA_FLAG_ID) (include/CodecUtils/IForwardTransform.h:56-126,
 *   src/CodecUtils/FastForward4x4ITImpl1.cpp:100-278).
 * icu_inv4x4 replaces IInverseTransform::InverseTransform + SetParameter +
 *   SetScale (include/CodecUtils/IInverseTransform.h:56-115,
 *   src/CodecUtils/FastInverse4x4ITImpl1.cpp:119-302); cols_first selects the
 *   FastInverse4x4On16x16ITImpl1 order (:120-360).  weights16 nullable (flat 16). */
int icu_fwd4x4(const int16_t* in, int16_t* out, long nblocks, int qp, int intra, int mode, int layout);
int icu_inv4x4(const int16_t* in, int16_t* out, long nblocks, int qp, int mode, int layout,
               int cols_first, const int* weights16);
int icu_fwd4x4_device(const int16_t* d_in, int16_t* d_out, long nblocks, int qp, int intra, int mode,
                      int layout, void* stream);
int icu_inv4x4_device(const int16_t* d_in, int16_t* d_out, long nblocks, int qp, int mode, int layout,
                      int cols_first, const int* d_weights16, void* stream);

/* Batched estimation of many streams of mixed size and range in one launch
 * sequence (no reference counterpart: the reference estimates one frame per
 * call).  cfgs[i] configures stream i; predictor must be 0 (zero). */
typedef struct icu_me_batch icu_me_batch;
int icu_me_batch_create(const icu_me_config* cfgs, int nstreams, icu_me_batch** out);
int icu_me_batch_destroy(icu_me_batch* b);
/* Total MB slots; stream i's MBs start at slot icu_me_batch_mb_base(b, i). */
int icu_me_batch_num_mbs(icu_me_batch* b);
int icu_me_batch_mb_base(icu_me_batch* b, int stream);
/* d_curs/d_refs: host arrays of nstreams device pointers (8-bit, stride W).
 * d_mv: 2 int16 per MB slot, d_dist: 1 int32 per slot (device). */
int icu_me_batch_estimate_device(icu_me_batch* b, const uint8_t* const* d_curs,
                                 const uint8_t* const* d_refs, int16_t* d_mv, int32_t* d_dist,
                                 void* stream);
/* Capture the whole batch (padding, the search launches with their in-kernel
 * decisions) into a HIP graph bound to these buffers, then replay it with
 * icu_me_batch_graph_launch.  With the queue plan (grid mode 4, the default
 * for a batch whose streams the one-MB-per-wave kernels take) the search is
 * one persistent launch over per-XCD macroblock queues. */
int icu_me_batch_capture(icu_me_batch* b, const uint8_t* const* d_curs, const uint8_t* const* d_refs,
                         int16_t* d_mv, int32_t* d_dist);
int icu_me_batch_graph_launch(icu_me_batch* b, void* stream);
/* Times the batch's launch plans on these buffers (each captured in a graph,
 * the least of three runs of `reps` replays) and keeps the fastest: grid
 * modes 0-2 (workgroup tiles for every stream), 3 (one launch per window
 * class), 5 (the same with the +-8 / +-16 classes on a second stream, beside
 * the others), 4 (the queue kernel at 6 or 8 waves per SIMD) and 6 (the
 * +-31/32 and +-63/64 classes in the queue kernel, the others as in 5).  *best_ms
 * (nullable): its time per replay.  It leaves no graph: capture again
 * afterwards. */
int icu_me_batch_autotune(icu_me_batch* b, const uint8_t* const* d_curs, const uint8_t* const* d_refs,
                          int16_t* d_mv, int32_t* d_dist, int reps, float* best_ms);
/* The batch's plan: grid mode (as icu_me_get_tiling reports it) and the queue
 * kernel's waves per SIMD.  set_plan pins one (a plan tuned earlier, say):
 * grid_mode 0..6 (3..6 need a stream the wave kernels take), queue_waves
 * 1..8; a graph captured before must be captured again. */
int icu_me_batch_get_plan(icu_me_batch* b, int* grid_mode, int* queue_waves);
int icu_me_batch_set_plan(icu_me_batch* b, int grid_mode, int queue_waves);

/* ------------------------------------------------- YUV motion compensation
 * Whole-frame YUV 4:2:0 compensation from a 16x16 quarter-pel vector list:
 * replaces MotionCompensatorH264ImplStd (ctor + Create + Compensate(pMotionList),
 * src/CodecUtils/MotionCompensatorH264ImplStd.cpp:76-240,323-473).  Luma per
 * MB: copy at integer vectors, else QuarterRead; chroma per 8x8 block at the
 * same vector in eighth pels: copy, else EighthRead's bilinear
 * (src/ImageUtils/OverlayMem2Dv2.cpp:1063-1114).  Reference samples outside
 * the picture repeat its edge. */
typedef struct icu_mc icu_mc;
int icu_mc_create(int width, int height, icu_mc** out);
int icu_mc_destroy(icu_mc* h);
/* In place on a host buffer laid out as the reference class's: int16 luma
 * W*H, then U and V (W/2)*(H/2), samples 0..255 (ICU_E_RANGE otherwise);
 * mv: 2 int16 per MB, raster order (VectorStructList SIMPLE2D). */
int icu_mc_compensate(icu_mc* h, int16_t* yuv, const int16_t* mv);
/* Device form on 8-bit planes (stride = plane width): the compensated
 * planes go to d_y/d_u/d_v, which may be the reference planes themselves. */
int icu_mc_compensate_device(icu_mc* h, const uint8_t* d_ref_y, const uint8_t* d_ref_u,
                             const uint8_t* d_ref_v, const int16_t* d_mv, uint8_t* d_y, uint8_t* d_u,
                             uint8_t* d_v, void* stream);

/* ------------------------------------------------------------------ CAVLC
 * Context-adaptive VLC of one block of quantised levels (H.264 9.2):
 * replaces CAVLCH264Impl::Encode / Decode with its VLC coders
 * (src/CodecUtils/CAVLCH264Impl.cpp:134-611; SetMode Mode2x2 / Mode4x4 /
 * Mode8x8, SetParameter NUM_TOT_NEIGHBOR_COEFF_ID = nc, DC_SKIP_FLAG_ID =
 * dc_skip) over an MSB-first bit stream (BitStreamWriterMSB /
 * BitStreamReaderMSB).  Blocks are 16 int16 levels in raster order (mode
 * ICU_CAVLC_4X4), 4 (ICU_CAVLC_2X2, chroma DC; nc -1 selects the 4:2:0
 * chroma DC coeff_token table, -2 the 4:2:2 one) or 64 (ICU_CAVLC_8X8,
 * scanned in the 8x8 zig-zag order but coded with the 4x4 tables, as the
 * reference's Mode8x8 does: a block codes when it has at most 15 levels
 * and at most 15 zeros below its last one; a (total_zeros, TotalCoeff)
 * pair without a 4x4 code fails with ICU_E_STATE after coeff_token and the
 * levels are written, as there; 16 levels or 16 such zeros, where the
 * reference reads past its tables, fail with ICU_E_STATE before anything
 * is written). */
#define ICU_CAVLC_2X2 0
#define ICU_CAVLC_4X4 1
#define ICU_CAVLC_8X8 2
/* Writes the block's code at *bit_pos of stream (stream_bits long; bits are
 * set or cleared, others untouched) and advances *bit_pos; stream NULL only
 * counts.  *nbits: bits written.  ICU_E_RANGE when the stream is full,
 * ICU_E_STATE when a symbol has no code; on either error the symbols written
 * before it stay, *bit_pos is advanced past them and *nbits counts them (the
 * reference's writer keeps them too).  A symbol is written whole or not at
 * all. */
int icu_cavlc_encode(const int16_t* coef, int mode, int nc, int dc_skip, uint8_t* stream, long stream_bits,
                     long* bit_pos, int* nbits);
/* Reads one block from *bit_pos (all mode-size outputs written; zeros
 * where no level), advances *bit_pos.  ICU_E_RANGE: the stream ended;
 * ICU_E_STATE: no such code. */
int icu_cavlc_decode(const uint8_t* stream, long stream_bits, long* bit_pos, int mode, int nc, int dc_skip,
                     int16_t* coef, int* nbits);
/* Batches on the device.  Encode nblocks blocks (block b with nC d_nc[b],
 * or nc for all when d_nc is NULL) into one MSB-first stream in block
 * order; d_offsets (nblocks + 1 int64) receives each block's first bit and,
 * last, the total, also returned in *total_bits.  Waits for the stream
 * length (one device-to-host read) before writing; ICU_E_RANGE when it
 * exceeds stream_bytes.  Decode takes such offsets; d_status (nullable)
 * receives each block's bits or a negative error. */
typedef struct icu_cavlc icu_cavlc;
int icu_cavlc_create(icu_cavlc** out);
int icu_cavlc_destroy(icu_cavlc* h);
int icu_cavlc_encode_device(icu_cavlc* h, const int16_t* d_coef, long nblocks, int mode, const int8_t* d_nc,
                            int nc, int dc_skip, uint8_t* d_stream, long stream_bytes, int64_t* d_offsets,
                            long long* total_bits, void* stream);
int icu_cavlc_decode_device(icu_cavlc* h, const uint8_t* d_stream, long stream_bytes, const int64_t* d_offsets,
                            long nblocks, int mode, const int8_t* d_nc, int nc, int dc_skip, int16_t* d_coef,
                            int32_t* d_status, void* stream);

/* ---------------------------------------------------------- UMHS estimator
 * MotionEstimatorH264ImplUMHS (include/CodecUtils/MotionEstimatorH264ImplUMHS.h:78-126,
 * src/CodecUtils/MotionEstimatorH264ImplUMHS.cpp:446-940): 16x16 SSD search
 * from the H.264 median predictor with distortion, the zero vector and the
 * previous frame's co-located vector; early-termination tests; unsymmetrical
 * cross, 5x5, uneven multi-hexagon, extended hexagon and diamond stages;
 * half / quarter refinement; a final test against the predictor.
 * motion_range: quarter pels, as the reference's constructor (8..256; full
 * pel window +-(motion_range/4 - 1) within the 16-pel border).  mode: the
 * reference's SetMode, 0 quarter, 1 half, 2 full pel.  The predictor is the
 * reference's H264MotionVectorPredictorImpl1 over the picture as one slice. */
typedef struct icu_umhs icu_umhs;
int icu_umhs_create(int width, int height, int motion_range, int mode, icu_umhs** out);
int icu_umhs_destroy(icu_umhs* h);
/* Host planes (int16 0..255, W*H).  prev_mv (2*nMB quarter-pel) and
 * prev_intra (nMB flags) are the previous frame's macroblocks (both NULL:
 * none, as if every one were intra; the reference class needs them).  A
 * previous vector must keep its block within the 16-pel border, as the
 * reference's own vectors do.  include (nullable) selects the MBs of *avg.
 * dist (nullable): per MB the distortion the predictor keeps. */
int icu_umhs_estimate(icu_umhs* h, const int16_t* cur, const int16_t* ref, const int16_t* prev_mv,
                      const uint8_t* prev_intra, const uint8_t* include, int16_t* mv, int32_t* dist, long* avg);
/* Device form: 8-bit planes (d_cur 16-byte aligned, stride W), asynchronous
 * on stream.  One wave per MB row runs the rows as a wavefront; calls that
 * overlap on one handle race.  icu_umhs_check waits for the device and
 * reports a wavefront that timed out (ICU_E_STATE; not expected). */
int icu_umhs_estimate_device(icu_umhs* h, const uint8_t* d_cur, const uint8_t* d_ref, const int16_t* d_prev_mv,
                             const uint8_t* d_prev_intra, int16_t* d_mv, int32_t* d_dist, void* stream);
int icu_umhs_check(icu_umhs* h);

/* --------------------------------------------------------- Cross estimator
 * MotionEstimatorH264ImplCross (include/CodecUtils/MotionEstimatorH264ImplCross.h:78-190,
 * src/CodecUtils/MotionEstimatorH264ImplCross.cpp:354-660): a cross search
 * whose arm halves from range/2 to 1, half / quarter refinement and
 * GetBestVector (lambda 0.9), every distortion the squared error over the
 * estimator's 64-point sampling path.  motion_range and mode as
 * icu_umhs_create; predictor 0 zero vector (MBs independent), 1 the H.264
 * median of the MBs decided before (H264MotionVectorPredictorImpl1; rows
 * run as a wavefront). */
typedef struct icu_cross icu_cross;
int icu_cross_create(int width, int height, int motion_range, int mode, int predictor, icu_cross** out);
int icu_cross_destroy(icu_cross* h);
/* Host planes (int16 0..255).  dist (nullable): per MB the path distortion
 * of the chosen vector; *avg over include (nullable) as the reference. */
int icu_cross_estimate(icu_cross* h, const int16_t* cur, const int16_t* ref, const uint8_t* include, int16_t* mv,
                       int32_t* dist, long* avg);
/* Device form (8-bit planes, stride W), asynchronous on stream; calls that
 * overlap on one handle race.  icu_cross_check as icu_umhs_check. */
int icu_cross_estimate_device(icu_cross* h, const uint8_t* d_cur, const uint8_t* d_ref, int16_t* d_mv,
                              int32_t* d_dist, void* stream);
int icu_cross_check(icu_cross* h);

/* ----------------------------------------------------------- FHS estimator
 * MotionEstimatorH264ImplFHS (include/CodecUtils/MotionEstimatorH264ImplFHS.h:78-214,
 * src/CodecUtils/MotionEstimatorH264ImplFHS.cpp:497-1790): from the H.264
 * median predictor with distortion, early exits, an uneven multi-hexagon,
 * one extended hexagon and one diamond over 80-point path distortions,
 * then HalfPelEstimate / QuarterPelEstimate (H.264 half and quarter
 * samples) and a final test against the predictor.  motion_range and mode
 * as icu_umhs_create.  The handle keeps the reference's half-sample cache
 * from call to call, so a handle codes one stream's frames in order, like
 * one reference instance; calls that overlap on one handle race. */
typedef struct icu_fhs icu_fhs;
int icu_fhs_create(int width, int height, int motion_range, int mode, icu_fhs** out);
int icu_fhs_destroy(icu_fhs* h);
int icu_fhs_estimate(icu_fhs* h, const int16_t* cur, const int16_t* ref, const uint8_t* include, int16_t* mv,
                     int32_t* dist, long* avg);
int icu_fhs_estimate_device(icu_fhs* h, const uint8_t* d_cur, const uint8_t* d_ref, int16_t* d_mv, int32_t* d_dist,
                            void* stream);
/* Waits for the device; ICU_E_STATE if a wavefront timed out (not
 * expected).  *unwritten_reads (nullable): MBs whose quarter candidates read
 * cache entries no call had written (the reference reads uninitialised
 * memory there; here they read 0). */
int icu_fhs_check(icu_fhs* h, int* unwritten_reads);

/* --------------------------------------------------- raw YUV and MV files
 * YuvRawFileHandler (include/CodecUtils/YuvRawFileHandler.h:53-98,
 * src/CodecUtils/YuvRawFileHandler.cpp:78-184 over RawFileHandlerBase,
 * src/CodecUtils/RawFileHandlerBase.cpp:82-180): a read handle maps the
 * file; each unit is one 4:2:0 frame from the current position, wrapping
 * to the start once the position reaches the last byte; a short last unit
 * reports how many samples it read (NV12: plus W*H/4, as the reference
 * counts) and keeps the previous unit's samples beyond them.  Types: the
 * handler's YUV4208P, YUV420NV12 (in only) and YUV42016P (out only). */
#define ICU_YUV420_8P 0
#define ICU_YUV420_NV12 1
#define ICU_YUV420_16P 16
typedef struct icu_yuv_file icu_yuv_file;
/* rw 0 read, 1 write (written on close, as the handler's delayed write). */
int icu_yuv_file_open(const char* path, int rw, int width, int height, int in_type, int out_type,
                      icu_yuv_file** out);
/* GetNextUnit: *unit points at the handle's buffer (int16 for 16P, uint8
 * for 8P) of W*H*3/2 samples, valid until the next call or close. */
int icu_yuv_file_next_unit(icu_yuv_file* f, const void** unit, int* unit_len);
/* PutNextUnit: appends bytes to the output stream. */
int icu_yuv_file_put_unit(icu_yuv_file* f, const void* unit, int bytes);
int icu_yuv_file_close(icu_yuv_file* f);
/* A raw frame already on the device (I420 or NV12 bytes) into the 8-bit
 * Y/U/V planes the estimator and compensator take (asynchronous). */
int icu_yuv_unpack_device(const uint8_t* d_raw, int in_type, int width, int height, uint8_t* d_y,
                          uint8_t* d_u, uint8_t* d_v, void* stream);
/* VectorStructList::Dump (src/CodecUtils/VectorStructList.cpp:242-270) of a
 * SIMPLE2D list: "mv x,mv y" then one "x,y" line per vector.  The title is
 * accepted and, as in the reference (which sets it before its table
 * exists), not written.  icu_mv_load reads such a file back into at most
 * max_n vectors (*n receives the count). */
int icu_mv_dump(const char* path, const char* title, const int16_t* mv, int n);
int icu_mv_load(const char* path, int16_t* mv, int max_n, int* n);

/* ---------------------------------------------------------------- serving
 * A request loop over many streams with continuous batching.  No reference
 * counterpart: the reference estimates and codes one frame per call on the
 * caller's thread (MotionEstimatorH264ImplFull::Estimate,
 * src/CodecUtils/MotionEstimatorH264ImplFull.cpp:330); the server runs the
 * same per-frame work (icu_me_estimate_device, then icu_code_frame_device at
 * the server's QP, inter, rows-first inverse) for many streams at once.
 *
 * Each stream keeps its reference frame (the previous frame's
 * reconstruction) resident in device memory; the first frame of a stream is
 * its own reference.  A submitted frame waits until its stream's previous
 * frame is done, fewer than max_inflight frames are on the device and its
 * transient buffers fit mem_budget; admitted frames run on `lanes` HIP
 * streams, so work of different streams overlaps and frames that arrive
 * while others run start as slots and memory free up.  Results depend only
 * on the stream's own frames, never on what else runs. */
typedef struct icu_server icu_server;
typedef struct {
  int lanes;           /* HIP streams work is spread over, 1..16 */
  int max_inflight;    /* frames on the device at once, >= 1 */
  long long mem_budget;/* device bytes for resident streams + frames in flight */
  int qp;              /* residual coding QP, 0..51 */
} icu_server_config;
typedef struct {
  long long tag;       /* the caller's tag from icu_server_submit */
  int stream, frame;   /* stream id, frame index within the stream */
  int nmb;             /* macroblocks (mv holds 2*nmb values) */
  double submit_s, done_s; /* icu_server_clock() at submit / at completion seen */
  unsigned long long levels_sum, recon_sum; /* checksums of the levels (int16
                        W*H, as bytes) and reconstruction (8-bit W*H): the sum
                        over bytes b[i] of (b[i] + 1) * ((i * 0x9E3779B1 mod 2^32) | 1),
                        mod 2^64 */
  const int16_t* mv;   /* quarter-pel vectors, valid until the next poll */
} icu_server_result;

int icu_server_create(const icu_server_config* cfg, icu_server** out);
/* Waits for frames in flight; queued frames are dropped. */
int icu_server_destroy(icu_server* s);
/* Opens a stream with estimator configuration cfg.  ICU_E_RANGE when its
 * resident state does not fit the remaining budget. */
int icu_server_open_stream(icu_server* s, const icu_me_config* cfg, int* stream_id);
/* Queues one 8-bit width*height luma frame (copied before return) and
 * admits what can start. */
int icu_server_submit(icu_server* s, int stream_id, const uint8_t* frame, long long tag);
/* Collects up to max_out finished frames (any order across streams, frame
 * order within a stream) and admits queued ones.  wait != 0: blocks until at
 * least one frame finishes when any is in flight.  Returns the number
 * collected (>= 0) or an error code. */
int icu_server_poll(icu_server* s, icu_server_result* out, int max_out, int wait);
/* The estimator handle of a stream (owned by the server, valid until
 * icu_server_destroy); calls on it race with the stream's frames in flight.
 * A poll that collects a median-predictor stream's frame whose wavefront
 * timed out (icu_me_check) returns ICU_E_STATE and drops that result. */
int icu_server_stream_estimator(icu_server* s, int stream_id, icu_me** out);
int icu_server_stats(icu_server* s, long long* bytes_in_use, long long* peak_bytes, int* queued,
                     int* inflight);
/* Seconds on the server's monotonic clock. */
double icu_server_clock(void);

/* ------------------------------------------ multi-reference, partitioned
 * Integer motion search over up to 4 reference frames with 16x16, 16x8,
 * 8x16 and 8x8 partitions, and the residual coding of its prediction.  No
 * reference counterpart: the reference's estimators search 16x16 blocks of
 * one reference.  Built from its pieces: 4x4 / 8x8 SADs as
 * OverlayMem2Dv2::Tad4x4 / Tad8x8 (src/ImageUtils/OverlayMem2Dv2.cpp:
 * 2555-2654), each MB's window as GetMotionRange with half-width range
 * (src/CodecUtils/MotionEstimatorH264ImplFull.cpp:1029-1081, exhaustive
 * form), shared by the MB's partitions.  Per partition p, reference r and
 * full-pel vector v: J = SAD_p(r, v) + lambda * (|vx| + |vy|); the least J
 * wins, ties to the lower r, then the earlier v in raster order.  The MB
 * mode is the least total J over its partitions, ties to fewer partitions. */
typedef struct icu_vbs icu_vbs;
typedef struct {
  int width, height; /* luma, multiples of 16 */
  int range;         /* full-pel half-width R, 1..32 */
  int nrefs;         /* reference frames, 1..4 */
  int lambda;        /* rate weight, >= 0, lambda * 2 * range < 65792 */
} icu_vbs_config;
int icu_vbs_create(const icu_vbs_config* cfg, icu_vbs** out);
int icu_vbs_destroy(icu_vbs* h);
/* d_refs: host array of nrefs device planes (8-bit, stride = width).
 * Outputs (device): d_mode per MB (0 16x16, 1 16x8, 2 8x16, 3 8x8);
 * d_mv8 2 int16 per 8x8 block (quarter pel, multiples of 4) and d_ref8 the
 * reference index per 8x8 block, both in 8x8-block raster order over the
 * picture; d_cost the chosen mode's total J per MB. */
int icu_vbs_estimate_device(icu_vbs* h, const uint8_t* d_cur, const uint8_t* const* d_refs,
                            uint8_t* d_mode, int16_t* d_mv8, uint8_t* d_ref8, int32_t* d_cost,
                            void* stream);
/* Residual coding at nqp (1..20) QPs of the prediction that copies each 8x8
 * block from reference d_ref8[b] (as padded by the last estimate) at
 * d_mv8[b] floored to full pels: levels / recon / pred as in
 * icu_code_frame_qps_device. */
int icu_vbs_code_device(icu_vbs* h, const uint8_t* d_cur, const int16_t* d_mv8, const uint8_t* d_ref8,
                        int nqp, const int* qps, int intra, int cols_first, int16_t* const* d_levels,
                        uint8_t* const* d_recon, int16_t* d_pred, void* stream);
/* Candidate rows per lane tile and workgroup LDS bytes of the search. */
int icu_vbs_get_plan(icu_vbs* h, int* tile_rows, int* lds_bytes);

/* Errors and a minimal device runtime for callers without their own. */
const char* icu_last_error(void);
int icu_device_info(char* name, int name_len, int* compute_units, int* clock_khz);
void* icu_malloc(size_t bytes);
int icu_free(void* d);
int icu_memcpy_htod(void* d, const void* h, size_t bytes, void* stream);
int icu_memcpy_dtoh(void* h, const void* d, size_t bytes, void* stream);
/* Asynchronous device-to-device copy on `stream`. */
int icu_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
int icu_memset(void* d, int value, size_t bytes, void* stream);
void* icu_stream_create(void);
int icu_stream_destroy(void* stream);
int icu_stream_sync(void* stream);
/* Timing events (device-scope release: after icu_event_record the host
 * reads device results through icu_memcpy_dtoh, not through the event). */
void* icu_event_create(void);
int icu_event_destroy(void* ev);
int icu_event_record(void* ev, void* stream);
int icu_event_elapsed_ms(void* start, void* stop, float* ms);

/* Correctly rounded sqrt of integers used by the final vector choice,
 * evaluated on the device for n values (test hook). */
int icu_sqrt_check_device(const int* d_m, double* d_out, int n, void* stream);

#ifdef __cplusplus
}
#endif
#endif
