// C entry points over the compiled reference classes, used only to write the
// golden fixtures and to time the reference CPU path.  TEST INFRASTRUCTURE:
// built by oracle/Makefile into oracle/_ref/ from the reference sources where
// they lie; nothing in the product links it.
//
// Classes driven (reference file:line):
//   MotionEstimatorH264ImplFull          src/CodecUtils/MotionEstimatorH264ImplFull.cpp:111-611
//   H264MotionVectorPredictorImpl1       include/CodecUtils/H264MotionVectorPredictorImpl1.h:54-132
//   MacroBlockH264::Initialise           src/CodecUtils/MacroBlockH264.cpp:318-400
//   FastForward4x4ITImpl1/ITImpl2/On16x16ITImpl1, FastInverse4x4ITImpl1/On16x16ITImpl1
//   OverlayMem2Dv2 / OverlayExtMem2Dv2   src/ImageUtils/OverlayMem2Dv2.cpp:704,1623,2655
//   MotionCompensatorH264ImplStd         src/CodecUtils/MotionCompensatorH264ImplStd.cpp:323-473
//   YuvRawFileHandler                    src/CodecUtils/YuvRawFileHandler.cpp:78-184
//   VectorStructList::Dump               src/CodecUtils/VectorStructList.cpp:242-270
//   CAVLCH264Impl + its VLC coders       src/CodecUtils/CAVLCH264Impl.cpp:134-577
//   MotionEstimatorH264ImplUMHS          src/CodecUtils/MotionEstimatorH264ImplUMHS.cpp:231-940
//   MotionEstimatorH264ImplCross         src/CodecUtils/MotionEstimatorH264ImplCross.cpp:137-660
//   MotionEstimatorH264ImplFHS           src/CodecUtils/MotionEstimatorH264ImplFHS.cpp:231-905
#include <math.h>  // the estimator header calls sqrt without including it

#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

#include "FastForward4x4ITImpl1.h"
#include "FastForward4x4ITImpl2.h"
#include "FastForward4x4On16x16ITImpl1.h"
#include "FastInverse4x4ITImpl1.h"
#include "FastInverse4x4On16x16ITImpl1.h"
#include "H264MotionVectorPredictorImpl1.h"
#include "MacroBlockH264.h"
#include "MotionCompensatorH264ImplStd.h"
#include "MotionEstimatorH264ImplFull.h"
#include "MotionEstimatorH264ImplUMHS.h"
#include "MotionEstimatorH264ImplCross.h"
#include "MotionEstimatorH264ImplFHS.h"
#include "OverlayExtMem2Dv2.h"
#include "OverlayMem2Dv2.h"
#include "VectorStructList.h"
#include "YuvRawFileHandler.h"
#include "BitStreamReaderMSB.h"
#include "BitStreamWriterMSB.h"
#include "CAVLCH264Impl.h"
#include "CoeffTokenH264VlcDecoder.h"
#include "CoeffTokenH264VlcEncoder.h"
#include "PrefixH264VlcDecoderImpl1.h"
#include "PrefixH264VlcEncoderImpl1.h"
#include "RunBeforeH264VlcDecoder.h"
#include "RunBeforeH264VlcEncoder.h"
#include "TotalZeros2x2H264VlcDecoder.h"
#include "TotalZeros2x2H264VlcEncoder.h"
#include "TotalZeros4x4H264VlcDecoder.h"
#include "TotalZeros4x4H264VlcEncoder.h"

namespace {

// A predictor that always predicts the zero vector (the GPU bench mode).
class ZeroPredictor : public IMotionVectorPredictor {
 public:
  int Create(void) override { return 1; }
  int Get16x16Prediction(void*, int, int* px, int* py) override {
    *px = 0;
    *py = 0;
    return 1;
  }
};

struct MeRun {
  int W, H, nmb;
  std::vector<short> cur, ref;
  std::vector<MacroBlockH264> mbs;
  std::vector<MacroBlockH264*> rows;
  IMotionVectorPredictor* pred = nullptr;
  MotionEstimatorH264ImplFull* me = nullptr;
  ~MeRun() {
    delete me;
    delete pred;
  }
};

MeRun* me_setup(int W, int H, int motionRange, int pred_mode, const short* cur,
                const short* ref, const unsigned char* include) {
  MeRun* r = new MeRun();
  r->W = W;
  r->H = H;
  r->nmb = (W / 16) * (H / 16);
  r->cur.assign(cur, cur + (long)W * H);
  r->ref.assign(ref, ref + (long)W * H);
  if (pred_mode) {
    int mbw = W / 16, mbh = H / 16;
    r->mbs.resize(r->nmb);
    r->rows.resize(mbh);
    for (int i = 0; i < mbh; i++) r->rows[i] = &r->mbs[i * mbw];
    MacroBlockH264::Initialise(mbh, mbw, 0, r->nmb - 1, 0, r->rows.data());
    r->pred = new H264MotionVectorPredictorImpl1(r->mbs.data());
  } else {
    r->pred = new ZeroPredictor();
  }
  if (include)
    r->me = new MotionEstimatorH264ImplFull(r->cur.data(), r->ref.data(), W, H, motionRange,
                                            r->pred, (void*)include);
  else
    r->me = new MotionEstimatorH264ImplFull(r->cur.data(), r->ref.data(), W, H, motionRange,
                                            r->pred);
  if (!r->me->Create()) {
    delete r;
    return nullptr;
  }
  return r;
}

}  // namespace

extern "C" {

// One Estimate() of the as-shipped class. include: bool per MB (nullable).
int ref_me_estimate(int W, int H, int motionRange, int pred_mode, const short* cur,
                    const short* ref, const unsigned char* include, short* mv_out,
                    long* avg_out) {
  static_assert(sizeof(bool) == 1, "bool include map");
  MeRun* r = me_setup(W, H, motionRange, pred_mode, cur, ref, include);
  if (!r) return -1;
  long avg = 0;
  VectorStructList* l = (VectorStructList*)r->me->Estimate(&avg);
  for (int i = 0; i < r->nmb; i++) {
    mv_out[2 * i] = (short)l->GetSimpleElement(i, 0);
    mv_out[2 * i + 1] = (short)l->GetSimpleElement(i, 1);
  }
  if (avg_out) *avg_out = avg;
  delete r;
  return 0;
}

// One Estimate() of MotionEstimatorH264ImplUMHS (the constructor with the
// previous frame's macroblocks; the shorter ones leave that pointer NULL,
// which Estimate dereferences) over an H264MotionVectorPredictorImpl1 on a
// fresh macroblock array.  prev_mv / prev_intra NULL: every previous MB
// intra.  dist_out: the distortion the predictor keeps per MB.
int ref_umhs_estimate(int W, int H, int motionRange, int mode, const short* cur, const short* ref,
                      const short* prev_mv, const unsigned char* prev_intra, const unsigned char* include,
                      short* mv_out, int* dist_out, long* avg_out, int reps, double* seconds) {
  const int mbw = W / 16, mbh = H / 16, nmb = mbw * mbh;
  std::vector<short> c(cur, cur + (long)W * H), r(ref, ref + (long)W * H);
  std::vector<MacroBlockH264> mbs(nmb), prev(nmb);
  std::vector<MacroBlockH264*> rows(mbh);
  for (int i = 0; i < mbh; i++) rows[i] = &mbs[(size_t)i * mbw];
  MacroBlockH264::Initialise(mbh, mbw, 0, nmb - 1, 0, rows.data());
  for (int k = 0; k < nmb; k++) {
    prev[k]._intraFlag = prev_mv ? (prev_intra[k] ? 1 : 0) : 1;
    prev[k]._mvX[MacroBlockH264::_16x16] = prev_mv ? prev_mv[2 * k] : 0;
    prev[k]._mvY[MacroBlockH264::_16x16] = prev_mv ? prev_mv[2 * k + 1] : 0;
  }
  H264MotionVectorPredictorImpl1 pred(mbs.data());
  MotionEstimatorH264ImplUMHS me(c.data(), r.data(), W, H, motionRange, &pred, (void*)include, prev.data());
  if (!me.Create()) return -1;
  me.SetMode(mode);
  long avg = 0;
  VectorStructList* l = nullptr;
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < (reps < 1 ? 1 : reps); i++) l = (VectorStructList*)me.Estimate(&avg);
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  for (int i = 0; i < nmb; i++) {
    if (mv_out) {
      mv_out[2 * i] = (short)l->GetSimpleElement(i, 0);
      mv_out[2 * i + 1] = (short)l->GetSimpleElement(i, 1);
    }
    if (dist_out) dist_out[i] = mbs[i]._distortion[0];
  }
  if (avg_out) *avg_out = avg;
  return 0;
}

// One Estimate() of MotionEstimatorH264ImplCross with the zero predictor
// (pred_mode 0) or H264MotionVectorPredictorImpl1 on a fresh macroblock
// array (1).  dist_out: the distortion GetBestVector chose, per MB (read
// back through a one-MB include map each, since the class reports only the
// average).
int ref_cross_estimate(int W, int H, int motionRange, int mode, int pred_mode, const short* cur, const short* ref,
                       const unsigned char* include, short* mv_out, long* avg_out, int reps, double* seconds) {
  const int mbw = W / 16, mbh = H / 16, nmb = mbw * mbh;
  std::vector<short> c(cur, cur + (long)W * H), r(ref, ref + (long)W * H);
  std::vector<MacroBlockH264> mbs;
  std::vector<MacroBlockH264*> rows(mbh);
  IMotionVectorPredictor* pred;
  if (pred_mode) {
    mbs.resize(nmb);
    for (int i = 0; i < mbh; i++) rows[i] = &mbs[(size_t)i * mbw];
    MacroBlockH264::Initialise(mbh, mbw, 0, nmb - 1, 0, rows.data());
    pred = new H264MotionVectorPredictorImpl1(mbs.data());
  } else {
    pred = new ZeroPredictor();
  }
  MotionEstimatorH264ImplCross* me =
      include ? new MotionEstimatorH264ImplCross(c.data(), r.data(), W, H, motionRange, pred, (void*)include)
              : new MotionEstimatorH264ImplCross(c.data(), r.data(), W, H, motionRange, pred);
  if (!me->Create()) {
    delete me;
    delete pred;
    return -1;
  }
  me->SetMode(mode);
  long avg = 0;
  VectorStructList* l = nullptr;
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < (reps < 1 ? 1 : reps); i++) l = (VectorStructList*)me->Estimate(&avg);
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  for (int i = 0; mv_out && i < nmb; i++) {
    mv_out[2 * i] = (short)l->GetSimpleElement(i, 0);
    mv_out[2 * i + 1] = (short)l->GetSimpleElement(i, 1);
  }
  if (avg_out) *avg_out = avg;
  delete me;
  delete pred;
  return 0;
}

// nframes Estimate() calls of one MotionEstimatorH264ImplFHS instance (its
// half-sample cache carries from call to call, as in a codec), frame f
// coding curs[f] against refs[f] (copied into the planes the class was
// built on).  Per frame: vectors, the distortion the predictor keeps per MB
// and the average over include.
int ref_fhs_chain(int W, int H, int motionRange, int mode, int nframes, const short* curs, const short* refs,
                  const unsigned char* include, short* mv_out, int* dist_out, long* avg_out, double* seconds) {
  const int mbw = W / 16, mbh = H / 16, nmb = mbw * mbh;
  const long np = (long)W * H;
  std::vector<short> c(np), r(np);
  std::vector<MacroBlockH264> mbs(nmb);
  std::vector<MacroBlockH264*> rows(mbh);
  for (int i = 0; i < mbh; i++) rows[i] = &mbs[(size_t)i * mbw];
  MacroBlockH264::Initialise(mbh, mbw, 0, nmb - 1, 0, rows.data());
  H264MotionVectorPredictorImpl1 pred(mbs.data());
  MotionEstimatorH264ImplFHS me(c.data(), r.data(), W, H, motionRange, &pred, (void*)include);
  if (!me.Create()) return -1;
  me.SetMode(mode);
  double t = 0;
  for (int f = 0; f < nframes; f++) {
    std::memcpy(c.data(), curs + f * np, sizeof(short) * np);
    std::memcpy(r.data(), refs + f * np, sizeof(short) * np);
    long avg = 0;
    auto t0 = std::chrono::steady_clock::now();
    VectorStructList* l = (VectorStructList*)me.Estimate(&avg);
    auto t1 = std::chrono::steady_clock::now();
    t += std::chrono::duration<double>(t1 - t0).count();
    for (int i = 0; i < nmb; i++) {
      mv_out[(long)f * 2 * nmb + 2 * i] = (short)l->GetSimpleElement(i, 0);
      mv_out[(long)f * 2 * nmb + 2 * i + 1] = (short)l->GetSimpleElement(i, 1);
      dist_out[(long)f * nmb + i] = mbs[i]._distortion[0];
    }
    avg_out[f] = avg;
  }
  if (seconds) *seconds = t;
  return 0;
}

// Times `reps` Estimate() calls on one constructed estimator; returns seconds.
double ref_me_time(int W, int H, int motionRange, int pred_mode, const short* cur,
                   const short* ref, int reps) {
  MeRun* r = me_setup(W, H, motionRange, pred_mode, cur, ref, nullptr);
  if (!r) return -1.0;
  long avg = 0;
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < reps; i++) r->me->Estimate(&avg);
  auto t1 = std::chrono::steady_clock::now();
  delete r;
  return std::chrono::duration<double>(t1 - t0).count();
}

// variant 0 = FastForward4x4ITImpl1, 1 = FastForward4x4ITImpl2 (n 4x4 blocks,
// 16 shorts each); 2 = FastForward4x4On16x16ITImpl1 (n 16x16 raster MBs).
void ref_fwd(int variant, short* data, long n, int qp, int intra, int mode) {
  IForwardTransform* t = variant == 0 ? (IForwardTransform*)new FastForward4x4ITImpl1()
                         : variant == 1 ? (IForwardTransform*)new FastForward4x4ITImpl2()
                                        : (IForwardTransform*)new FastForward4x4On16x16ITImpl1();
  t->SetParameter(IForwardTransform::INTRA_FLAG_ID, intra);
  t->SetParameter(IForwardTransform::QUANT_ID, qp);
  t->SetMode(mode);
  const long step = variant == 2 ? 256 : 16;
  for (long i = 0; i < n; i++) t->Transform(data + step * i);
  delete t;
}

// variant 0 = FastInverse4x4ITImpl1 (4x4 blocks), 1 = FastInverse4x4On16x16ITImpl1
// (16x16 raster MBs). weight16: optional SetScale array.
void ref_inv(int variant, short* data, long n, int qp, int mode, const int* weight16) {
  IInverseTransform* t = variant == 0 ? (IInverseTransform*)new FastInverse4x4ITImpl1()
                                      : (IInverseTransform*)new FastInverse4x4On16x16ITImpl1();
  if (weight16) t->SetScale((void*)weight16);
  t->SetParameter(IInverseTransform::QUANT_ID, qp);
  t->SetMode(mode);
  const long step = variant == 1 ? 256 : 16;
  for (long i = 0; i < n; i++) t->InverseTransform(data + step * i);
  delete t;
}

// OverlayMem2Dv2::QuarterRead of a 16x16 block at full-pel (x,y) of a
// reference padded by 19 (OverlayExtMem2Dv2), quarter offsets qx,qy in -3..3.
void ref_quarter_read(const short* ref, int W, int H, int x, int y, int qx, int qy,
                      short* out256) {
  short* ext = nullptr;
  OverlayExtMem2Dv2::ExtendBoundary((void*)ref, W, H, 19, 19, (void**)&ext);
  OverlayExtMem2Dv2 ov(ext, W + 38, H + 38, 16, 16, 19, 19);
  ov.SetOrigin(x, y);
  short blk[256];
  OverlayMem2Dv2 dst(blk, 16, 16, 16, 16);
  ov.QuarterRead(dst, qx, qy);
  std::memcpy(out256, blk, sizeof(blk));
  delete[] ext;
}

// Tsd16x16 (metric 0) or Tad16x16 (metric 1) between two 16x16 blocks.
int ref_block_dist(const short* a256, const short* b256, int metric) {
  OverlayMem2Dv2 a((void*)a256, 16, 16, 16, 16), b((void*)b256, 16, 16, 16, 16);
  return metric ? a.Tad16x16(b) : a.Tsd16x16(b);
}

// Tsd16x16LessThan / Tad16x16LessThan with an explicit bound.
int ref_block_dist_lt(const short* a256, const short* b256, int metric, int min) {
  OverlayMem2Dv2 a((void*)a256, 16, 16, 16, 16), b((void*)b256, 16, 16, 16, 16);
  return metric ? a.Tad16x16LessThan(b, min) : a.Tsd16x16LessThan(b, min);
}

// Luma part of MotionCompensatorH264ImplStd::Compensate(list): the
// prediction is written into (a copy of) the reference luma plane.
int ref_mc_luma(const short* ref, int W, int H, int range, const short* mv, short* pred_out) {
  long lum = (long)W * H, chr = lum / 4;
  std::vector<short> yuv(lum + 2 * chr, 128);
  std::memcpy(yuv.data(), ref, lum * sizeof(short));
  MotionCompensatorH264ImplStd mc(range);
  if (!mc.Create(yuv.data(), W, H, 16, 16)) return -1;
  int nmb = (W / 16) * (H / 16);
  VectorStructList l(VectorStructList::SIMPLE2D);
  l.SetLength(nmb);
  for (int i = 0; i < nmb; i++) {
    l.SetSimpleElement(i, 0, mv[2 * i]);
    l.SetSimpleElement(i, 1, mv[2 * i + 1]);
  }
  mc.Compensate(&l);
  std::memcpy(pred_out, yuv.data(), lum * sizeof(short));
  return 0;
}

// Whole-frame YUV 4:2:0 compensation in place (luma then both chroma planes),
// MotionCompensatorH264ImplStd::Compensate(pMotionList).
int ref_mc_yuv(short* yuv, int W, int H, int range, const short* mv) {
  MotionCompensatorH264ImplStd mc(range);
  if (!mc.Create(yuv, W, H, 16, 16)) return -1;
  int nmb = (W / 16) * (H / 16);
  VectorStructList l(VectorStructList::SIMPLE2D);
  l.SetLength(nmb);
  for (int i = 0; i < nmb; i++) {
    l.SetSimpleElement(i, 0, mv[2 * i]);
    l.SetSimpleElement(i, 1, mv[2 * i + 1]);
  }
  mc.Compensate(&l);
  return 0;
}

// Reads `units` units with YuvRawFileHandler::GetNextUnit; out receives
// units x (frame elements) samples as int (16-bit or 8-bit outputs widened),
// lens the unit lengths.
int ref_yuv_read(const char* path, int W, int H, int in_type, int out_type, int units, int* out, int* lens) {
  YuvRawFileHandler f;
  f.SetWidth(W);
  f.SetHeight(H);
  f.SetType(in_type, out_type);
  if (!f.Open(path, 0)) return -1;
  const int n = W * H * 3 / 2;
  for (int u = 0; u < units; u++) {
    int len = 0;
    void* p = f.GetNextUnit(&len);
    lens[u] = len;
    for (int i = 0; i < n; i++)
      out[(long)u * n + i] = out_type == YuvRawFileHandler::YUV42016P ? ((short*)p)[i] : ((unsigned char*)p)[i];
  }
  f.Close();
  return 0;
}

int ref_mv_dump(const char* path, const char* title, const short* mv, int n) {
  VectorStructList l(VectorStructList::SIMPLE2D);
  l.SetLength(n);
  for (int i = 0; i < n; i++) {
    l.SetSimpleElement(i, 0, mv[2 * i]);
    l.SetSimpleElement(i, 1, mv[2 * i + 1]);
  }
  l.Dump((char*)path, title);
  return 0;
}

// CAVLC of n blocks (4 or 16 shorts each) into one MSB-first stream, block
// b with nC nc[b]; bits[b] receives Encode's return value.  Then the stream
// is decoded block by block into dec (dec_bits: Decode's return value).
int ref_cavlc(const short* blocks, int n, int mode, const int* nc, int dc_skip, unsigned char* out,
              int out_bytes, int* bits, short* dec, int* dec_bits) {
  PrefixH264VlcEncoderImpl1 pe;
  PrefixH264VlcDecoderImpl1 pd;
  CoeffTokenH264VlcEncoder ce;
  CoeffTokenH264VlcDecoder cd;
  TotalZeros4x4H264VlcEncoder t4e;
  TotalZeros4x4H264VlcDecoder t4d;
  TotalZeros2x2H264VlcEncoder t2e;
  TotalZeros2x2H264VlcDecoder t2d;
  RunBeforeH264VlcEncoder re;
  RunBeforeH264VlcDecoder rd;
  CAVLCH264Impl c;
  c.SetTokenCoeffVlcEncoder(&ce);
  c.SetTokenCoeffVlcDecoder(&cd);
  c.SetPrefixVlcEncoder(&pe);
  c.SetPrefixVlcDecoder(&pd);
  c.SetRunBeforeVlcEncoder(&re);
  c.SetRunBeforeVlcDecoder(&rd);
  if (mode == CAVLCH264Impl::Mode2x2) {
    c.SetTotalZerosVlcEncoder(&t2e);
    c.SetTotalZerosVlcDecoder(&t2d);
  } else {
    c.SetTotalZerosVlcEncoder(&t4e);
    c.SetTotalZerosVlcDecoder(&t4d);
  }
  c.SetMode(mode);
  const int size = mode == CAVLCH264Impl::Mode2x2 ? 4 : 16;
  std::memset(out, 0, out_bytes);
  BitStreamWriterMSB w;
  w.SetStream(out, out_bytes * 8);
  for (int b = 0; b < n; b++) {
    c.SetParameter(CAVLCH264Impl::NUM_TOT_NEIGHBOR_COEFF_ID, nc[b]);
    c.SetParameter(CAVLCH264Impl::DC_SKIP_FLAG_ID, dc_skip);
    // the coder casts its void* back to the interface: pass the interface
    // subobject (the writer derives from more than one base)
    bits[b] = c.Encode((void*)(blocks + (long)b * size), (void*)static_cast<IBitStreamWriter*>(&w));
  }
  BitStreamReaderMSB r;
  r.SetStream(out, out_bytes * 8);
  for (int b = 0; b < n; b++) {
    c.SetParameter(CAVLCH264Impl::NUM_TOT_NEIGHBOR_COEFF_ID, nc[b]);
    c.SetParameter(CAVLCH264Impl::DC_SKIP_FLAG_ID, dc_skip);
    dec_bits[b] = c.Decode((void*)static_cast<IBitStreamReader*>(&r), (void*)(dec + (long)b * size));
  }
  return 0;
}

// CAVLC of n blocks of `size` shorts (mode 0: 4, 1: 16, 2: 64), each into
// its own stream of block_bytes bytes (out + b * block_bytes), so that a
// block whose Encode fails leaves no trace in the next one.  bits[b]:
// Encode's return value; used[b]: the bits the writer holds afterwards (a
// failed Encode keeps the symbols it wrote before the failure).  Blocks
// encoded without error are decoded back from their stream into dec
// (dec_bits: Decode's return value).
int ref_cavlc_each(const short* blocks, int n, int mode, const int* nc, int dc_skip, unsigned char* out,
                   int block_bytes, int* bits, int* used, short* dec, int* dec_bits) {
  PrefixH264VlcEncoderImpl1 pe;
  PrefixH264VlcDecoderImpl1 pd;
  CoeffTokenH264VlcEncoder ce;
  CoeffTokenH264VlcDecoder cd;
  TotalZeros4x4H264VlcEncoder t4e;
  TotalZeros4x4H264VlcDecoder t4d;
  TotalZeros2x2H264VlcEncoder t2e;
  TotalZeros2x2H264VlcDecoder t2d;
  RunBeforeH264VlcEncoder re;
  RunBeforeH264VlcDecoder rd;
  CAVLCH264Impl c;
  c.SetTokenCoeffVlcEncoder(&ce);
  c.SetTokenCoeffVlcDecoder(&cd);
  c.SetPrefixVlcEncoder(&pe);
  c.SetPrefixVlcDecoder(&pd);
  c.SetRunBeforeVlcEncoder(&re);
  c.SetRunBeforeVlcDecoder(&rd);
  if (mode == CAVLCH264Impl::Mode2x2) {
    c.SetTotalZerosVlcEncoder(&t2e);
    c.SetTotalZerosVlcDecoder(&t2d);
  } else {
    c.SetTotalZerosVlcEncoder(&t4e);
    c.SetTotalZerosVlcDecoder(&t4d);
  }
  c.SetMode(mode);
  const int size = mode == CAVLCH264Impl::Mode2x2 ? 4 : (mode == CAVLCH264Impl::Mode8x8 ? 64 : 16);
  for (int b = 0; b < n; b++) {
    unsigned char* s = out + (long)b * block_bytes;
    std::memset(s, 0, block_bytes);
    c.SetParameter(CAVLCH264Impl::NUM_TOT_NEIGHBOR_COEFF_ID, nc[b]);
    c.SetParameter(CAVLCH264Impl::DC_SKIP_FLAG_ID, dc_skip);
    BitStreamWriterMSB w;
    w.SetStream(s, block_bytes * 8);
    bits[b] = c.Encode((void*)(blocks + (long)b * size), (void*)static_cast<IBitStreamWriter*>(&w));
    used[b] = block_bytes * 8 - w.GetStreamBitsRemaining();
    short* d = dec + (long)b * size;
    std::memset(d, 0, sizeof(short) * size);
    dec_bits[b] = 0;
    if (bits[b] > 0) {
      BitStreamReaderMSB r;
      r.SetStream(s, block_bytes * 8);
      dec_bits[b] = c.Decode((void*)static_cast<IBitStreamReader*>(&r), (void*)d);
    }
  }
  return 0;
}

}  // extern "C"
