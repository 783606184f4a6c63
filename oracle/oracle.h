/*
 * CPU oracle for the motion-estimation and 4x4 transform hot path.
 * TEST INFRASTRUCTURE ONLY (see me_oracle.c / it_oracle.c headers).
 */
#ifndef ICU_ORACLE_H
#define ICU_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int width, height; /* luma, multiples of 16 */
  int range;         /* full-pel search range R (reference: motionRange/4) */
  int metric;        /* 0 SSD (as shipped), 1 SAD */
  int window;        /* 0 reference rule +-(R-1), 1 exhaustive +-R */
  int filter;        /* 1 running-best candidate filter (as shipped), 0 off */
  int subpel;        /* 0 integer, 1 half, 2 half+quarter (as shipped) */
  int pred;          /* 0 zero predictor, 1 H.264 median (as shipped) */
  int bestvec;       /* 1 GetBestVector (as shipped), 0 off */
} orc_me_cfg;

/* mv_out: 2*nMB int16 quarter-pel (x,y) raster order.  dist_out: nMB
 * distortion of the chosen vector (nullable).  int_out: 3*nMB integer-stage
 * (mx, my, minDiff) (nullable).  include: nMB flags (nullable). */
int orc_me_estimate(const orc_me_cfg *cfg, const int16_t *cur, const int16_t *ref,
                    const uint8_t *include, int16_t *mv_out, int32_t *dist_out,
                    int32_t *int_out, long *avg_out);
int orc_me_estimate_subset(const orc_me_cfg *cfg, const int16_t *cur, const int16_t *ref,
                           const uint8_t *include, int step, int16_t *mv_out, int32_t *dist_out,
                           int32_t *int_out, long *avg_out);
int orc_me_estimate_phase(const orc_me_cfg *cfg, const int16_t *cur, const int16_t *ref,
                          const uint8_t *include, int step, int phase, int16_t *mv_out,
                          int32_t *dist_out, int32_t *int_out, long *avg_out);
int orc_qpel_sample(const int16_t *ref, int W, int H, int x, int y, int fx, int fy);
int orc_block_dist(const int16_t *a256, const int16_t *b256, int metric);
void orc_quarter_read(const int16_t *ref, int W, int H, int x, int y, int qx, int qy,
                      int *blk);
void orc_median_pred(const int16_t *mv, int mbw, int col, int row, int *px, int *py);

/* MotionEstimatorH264ImplUMHS (me_oracle.c).  motion_range in quarter pels
 * as the reference's constructor; mode 0 quarter, 1 half, 2 full pel.
 * prev_mv / prev_intra: the previous frame's co-located 16x16 vectors and
 * intra flags (NULL: no previous-frame candidate).  dist_out: the chosen
 * vector's SSD, which the predictor also keeps. */
int orc_umhs_estimate(const int16_t *cur, const int16_t *ref, int W, int H, int motion_range, int mode,
                      const int16_t *prev_mv, const uint8_t *prev_intra, const uint8_t *include,
                      int16_t *mv_out, int32_t *dist_out, long *avg_out);
void orc_umhs_stats(long *out13);
/* MotionEstimatorH264ImplCross (me_oracle.c): pred 0 zero predictor, 1 the
 * H.264 median; distortions are the estimator's 64-point path sums. */
void orc_cross_stats(long *out6);
/* MotionEstimatorH264ImplFHS (me_oracle.c): median predictor with
 * distortion; cache / cache_written (3*18*18 each) carry the estimator's
 * half-sample cache from call to call (zeros before the first). */
int orc_fhs_estimate(const int16_t *cur, const int16_t *ref, int W, int H, int qrange, int mode,
                     const uint8_t *include, int32_t *cache, uint8_t *cache_written, int16_t *mv_out,
                     int32_t *dist_out, long *avg_out);
void orc_fhs_stats(long *out8);
int orc_cross_estimate(const int16_t *cur, const int16_t *ref, int W, int H, int qrange, int mode, int pred,
                       const uint8_t *include, int16_t *mv_out, int32_t *dist_out, long *avg_out);
void orc_umhs_pred(const int16_t *mv, const int32_t *dist, int mbw, int col, int row, int *px, int *py,
                   int *pd);

/* 4x4 transforms.  mode: 0 IT+Q, 1 IT only, 2 Q only. */
void orc_fwd4x4(int16_t *blk, int qp, int intra, int mode);
void orc_inv4x4(int16_t *blk, int qp, int mode, const int *weight16);
void orc_inv4x4_colsfirst(int16_t *blk, int qp, int mode, const int *weight16);
/* Batched over n contiguous 4x4 blocks (16 int16 each). */
void orc_fwd4x4_n(int16_t *blks, long n, int qp, int intra, int mode);
void orc_inv4x4_n(int16_t *blks, long n, int qp, int mode, int cols_first);
/* Whole 16x16 raster macroblocks (FastForward4x4On16x16ITImpl1 layout). */
void orc_fwd_mb_n(int16_t *mbs, long n, int qp, int intra, int mode);
void orc_inv_mb_n(int16_t *mbs, long n, int qp, int mode, int cols_first);

/* Luma motion compensation of a whole frame from a quarter-pel MV list
 * (MotionCompensatorH264ImplStd::Compensate, luma part). pred_out W*H. */
void orc_mc_luma(const int16_t *ref, int W, int H, const int16_t *mv, int16_t *pred_out);

/* Residual -> fwd IT+Q -> levels; levels -> dequant+inverse -> recon with
 * clip, per 4x4 of each 16x16 MB.  levels_out W*H in MB raster layout. */
void orc_code_residual(const int16_t *cur, const int16_t *pred, int W, int H, int qp,
                       int intra, int cols_first, int16_t *levels_out, int16_t *recon_out);

/* Chroma compensation of one 4:2:0 chroma plane (Wc x Hc) from the per-MB
 * quarter-pel luma vectors (mc_oracle.c). */
void orc_mc_chroma(const uint8_t *ref, int Wc, int Hc, const int16_t *mv, uint8_t *out);

/* Multi-reference variable-block-size integer search (vbs_oracle.c).
 * refs: nrefs contiguous 8-bit W*H planes.  mode_out: nMB (0 16x16, 1 16x8,
 * 2 8x16, 3 8x8); mv8_out: 2 int16 quarter-pel per 8x8 block, 8x8-block
 * raster over the picture; ref8_out: reference index per 8x8 block;
 * cost_out: the chosen mode's total J per MB. */
int orc_vbs_estimate(const uint8_t *cur, const uint8_t *refs, int nrefs, int W, int H, int R,
                     int lambda, uint8_t *mode_out, int16_t *mv8_out, uint8_t *ref8_out,
                     int32_t *cost_out);
void orc_vbs_pred(const uint8_t *refs, int W, int H, const int16_t *mv8, const uint8_t *ref8,
                  int16_t *pred_out);

#ifdef __cplusplus
}
#endif
#endif
