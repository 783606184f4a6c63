// Drives the C++ drop-in classes the way codec code drives the reference
// classes (construct, Create, Estimate, read the VectorStructList; transform
// objects with SetParameter / SetMode / Transform in place).  Inputs and
// outputs are raw files; tests/test_gpu_dropin.py compares them.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "codecutils_hip.hpp"

using namespace icu_codecutils;

template <typename T>
static std::vector<T> load(const char* path, size_t n) {
  std::vector<T> v(n);
  FILE* f = std::fopen(path, "rb");
  if (!f || std::fread(v.data(), sizeof(T), n, f) != n) {
    std::fprintf(stderr, "cannot read %s\n", path);
    std::exit(2);
  }
  std::fclose(f);
  return v;
}

template <typename T>
static void save(const char* path, const T* p, size_t n) {
  FILE* f = std::fopen(path, "wb");
  if (!f || std::fwrite(p, sizeof(T), n, f) != n) std::exit(3);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 2) return 1;
  const std::string cmd = argv[1];
  if (cmd == "me" && argc >= 9) {
    const int W = std::atoi(argv[2]), H = std::atoi(argv[3]), range = std::atoi(argv[4]);
    const int median = std::atoi(argv[5]);
    auto cur = load<short>(argv[6], (size_t)W * H);
    auto ref = load<short>(argv[7], (size_t)W * H);
    const int nmb = (W / 16) * (H / 16);
    std::vector<unsigned char> inc;
    if (argc >= 10) inc = load<unsigned char>(argv[9], nmb);
    H264MotionVectorPredictorImpl1 med(W / 16, H / 16);
    ZeroMotionVectorPredictor zero;
    IMotionVectorPredictor* pred = median ? (IMotionVectorPredictor*)&med : (IMotionVectorPredictor*)&zero;
    MotionEstimatorH264ImplFull me(cur.data(), ref.data(), W, H, range, pred,
                                   inc.empty() ? nullptr : (void*)inc.data());
    if (!me.Create()) {
      std::fprintf(stderr, "Create failed: %s\n", me.LastError());
      return 4;
    }
    long avg = -1;
    VectorStructList* l = (VectorStructList*)me.Estimate(&avg);
    if (!l) {
      std::fprintf(stderr, "Estimate failed: %s\n", me.LastError());
      return 5;
    }
    std::vector<short> mv(2 * nmb);
    for (int i = 0; i < nmb; i++) {
      mv[2 * i] = (short)l->GetSimpleElement(i, 0);
      mv[2 * i + 1] = (short)l->GetSimpleElement(i, 1);
    }
    save(argv[8], mv.data(), mv.size());
    // the predictor now holds the decided vectors, as after the reference's Estimate
    int px = 0, py = 0;
    pred->Get16x16Prediction(nullptr, nmb - 1, &px, &py);
    std::printf("avg %ld pred_last %d %d\n", avg, px, py);
    return 0;
  }
  if ((cmd == "fwd" || cmd == "inv") && argc >= 9) {
    const int layout = std::atoi(argv[2]), qp = std::atoi(argv[3]), a = std::atoi(argv[4]);
    const int mode = std::atoi(argv[5]);
    const long calls = std::atol(argv[8]);
    const size_t n = (size_t)calls * (layout ? 256 : 16);
    auto x = load<short>(argv[6], n);
    if (cmd == "fwd") {
      if (layout) {
        FastForward4x4On16x16ITImpl1 t;
        t.SetParameter(IForwardTransform::INTRA_FLAG_ID, a);
        t.SetParameter(IForwardTransform::QUANT_ID, qp);
        t.SetMode(mode);
        t.TransformBlocks(x.data(), calls);
      } else {
        FastForward4x4ITImpl1 t;
        t.SetParameter(IForwardTransform::INTRA_FLAG_ID, a);
        t.SetParameter(IForwardTransform::QUANT_ID, qp);
        t.SetMode(mode);
        // one Transform call per block, as codec code calls the reference
        for (long i = 0; i < (calls < 4 ? calls : 4); i++) t.Transform(x.data() + 16 * i);
        if (calls > 4) t.TransformBlocks(x.data() + 64, calls - 4);
      }
    } else {
      if (layout) {
        FastInverse4x4On16x16ITImpl1 t;
        t.SetParameter(IInverseTransform::QUANT_ID, qp);
        t.SetMode(mode);
        t.InverseTransformBlocks(x.data(), calls);
      } else {
        FastInverse4x4ITImpl1 t;
        t.SetParameter(IInverseTransform::QUANT_ID, qp);
        t.SetMode(mode);
        t.InverseTransformBlocks(x.data(), calls);
      }
    }
    save(argv[7], x.data(), n);
    return 0;
  }
  if (cmd == "mc" && argc >= 6) {
    // mc W H yuv.raw mv.raw: compensates the YUV buffer in place, as
    // MotionCompensatorH264ImplStd::Compensate(pMotionList) does
    const int W = std::atoi(argv[2]), H = std::atoi(argv[3]);
    const size_t n = (size_t)W * H * 3 / 2;
    auto yuv = load<short>(argv[4], n);
    const int nmb = (W / 16) * (H / 16);
    auto mv = load<short>(argv[5], 2 * (size_t)nmb);
    VectorStructList l(VectorStructList::SIMPLE2D);
    l.SetLength(nmb);
    for (int i = 0; i < nmb; i++) {
      l.SetSimpleElement(i, 0, mv[2 * i]);
      l.SetSimpleElement(i, 1, mv[2 * i + 1]);
    }
    MotionCompensatorH264ImplStd mc(32);
    if (!mc.Create(yuv.data(), W, H, 16, 16)) {
      std::fprintf(stderr, "Create failed: %s\n", mc.LastError());
      return 4;
    }
    mc.Compensate(&l);
    save(argv[4], yuv.data(), n);
    return 0;
  }
  if (cmd == "cavlc" && argc >= 9) {
    // cavlc mode dc n blocks.raw nc.raw stream.raw out.raw: every block
    // through CAVLCH264Impl::Encode into one BitStreamWriterMSB stream
    // (SetParameter(NUM_TOT_NEIGHBOR_COEFF_ID) per block), then Decode from
    // a BitStreamReaderMSB; out.raw holds per block the encoded bits, the
    // decoded bits, NUM_TOT_COEFF after each and the decoded levels.
    const int mode = std::atoi(argv[2]), dc = std::atoi(argv[3]), n = std::atoi(argv[4]);
    const int size = mode == IContextAwareRunLevelCodec::Mode2x2 ? 4 : 16;
    auto blocks = load<short>(argv[5], (size_t)n * size);
    auto nc = load<signed char>(argv[6], (size_t)n);
    std::vector<unsigned char> stream((size_t)n * 128 + 16, 0xA5);
    CAVLCH264Impl codec;
    codec.SetMode(mode);
    codec.SetParameter(IContextAwareRunLevelCodec::DC_SKIP_FLAG_ID, dc);
    BitStreamWriterMSB w;
    w.SetStream(stream.data(), (int)stream.size() * 8);
    std::vector<int> out;
    for (int i = 0; i < n; i++) {
      codec.SetParameter(IContextAwareRunLevelCodec::NUM_TOT_NEIGHBOR_COEFF_ID, nc[i]);
      out.push_back(codec.Encode(&blocks[(size_t)i * size], &w));
      out.push_back(codec.GetParameter(IContextAwareRunLevelCodec::NUM_TOT_COEFF_ID));
    }
    const int total = (int)(stream.size() * 8) - w.GetStreamBitsRemaining();
    save(argv[7], stream.data(), (size_t)(total + 7) / 8);
    BitStreamReaderMSB r;
    r.SetStream(stream.data(), total);
    std::vector<short> dec(size);
    for (int i = 0; i < n; i++) {
      codec.SetParameter(IContextAwareRunLevelCodec::NUM_TOT_NEIGHBOR_COEFF_ID, nc[i]);
      out.push_back(codec.Decode(&r, dec.data()));
      out.push_back(codec.GetParameter(IContextAwareRunLevelCodec::NUM_TOT_COEFF_ID));
      for (int k = 0; k < size; k++) out.push_back(dec[k]);
    }
    // the cursor round trip of the reference's position encoding
    w.Reset();
    w.Write(13, 0x1abc);
    const int p = w.GetStreamBitPos();
    w.Write(5, 0);
    out.push_back(w.Seek(p));
    out.push_back(w.GetStreamBitsRemaining());
    save(argv[8], out.data(), out.size());
    return 0;
  }
  if (cmd == "fast" && argc >= 15) {
    // fast umhs|cross|fhs W H range mode pred nframes curs.raw refs.raw
    //      include.raw prev_mv.raw|- prev_intra.raw|- out.raw:
    // frames through one estimator object built on two planes (each frame
    // copied in before Estimate, as a codec refills its buffers); out.raw
    // holds per frame the vectors, then every frame's average distortion.
    const std::string kind = argv[2];
    const int W = std::atoi(argv[3]), H = std::atoi(argv[4]), range = std::atoi(argv[5]);
    const int mode = std::atoi(argv[6]), pred_median = std::atoi(argv[7]), nf = std::atoi(argv[8]);
    const size_t np = (size_t)W * H;
    const int nmb = (W / 16) * (H / 16);
    auto curs = load<short>(argv[9], np * nf);
    auto refs = load<short>(argv[10], np * nf);
    auto inc8 = load<unsigned char>(argv[11], nmb);
    std::vector<short> pmv;
    std::vector<unsigned char> pin;
    if (std::string(argv[12]) != "-") {
      pmv = load<short>(argv[12], 2 * (size_t)nmb);
      pin = load<unsigned char>(argv[13], nmb);
    }
    // the include map as the reference's bool array (one byte, 0 or 1, each)
    std::vector<unsigned char> incb(nmb);
    for (int i = 0; i < nmb; i++) incb[i] = inc8[i] != 0;
    void* inc = incb.data();
    std::vector<short> cur(np), ref(np);
    H264MotionVectorPredictorImpl1 med(W / 16, H / 16);
    ZeroMotionVectorPredictor zero;
    IMotionVectorPredictor* pred = pred_median ? (IMotionVectorPredictor*)&med : (IMotionVectorPredictor*)&zero;
    IMotionEstimator* me;
    if (kind == "umhs")
      me = new MotionEstimatorH264ImplUMHS(cur.data(), ref.data(), W, H, range, pred, inc,
                                           pmv.empty() ? nullptr : pmv.data(), pin.empty() ? nullptr : pin.data());
    else if (kind == "cross")
      me = new MotionEstimatorH264ImplCross(cur.data(), ref.data(), W, H, range, pred, inc);
    else
      me = new MotionEstimatorH264ImplFHS(cur.data(), ref.data(), W, H, range, pred, inc);
    if (!me->Create()) return 4;
    me->SetMode(mode);
    std::vector<short> out;
    std::vector<long> avgs;
    for (int f = 0; f < nf; f++) {
      std::copy(curs.begin() + f * np, curs.begin() + (f + 1) * np, cur.begin());
      std::copy(refs.begin() + f * np, refs.begin() + (f + 1) * np, ref.begin());
      long avg = 0;
      VectorStructList* l = (VectorStructList*)me->Estimate(&avg);
      if (!l) {
        std::fprintf(stderr, "Estimate failed: %s\n", icu_last_error());
        return 5;
      }
      for (int i = 0; i < nmb; i++) {
        out.push_back((short)l->GetSimpleElement(i, 0));
        out.push_back((short)l->GetSimpleElement(i, 1));
      }
      avgs.push_back(avg);
    }
    delete me;
    FILE* fo = std::fopen(argv[14], "wb");
    if (!fo) return 3;
    std::fwrite(out.data(), sizeof(short), out.size(), fo);
    for (long a : avgs) {
      const long long v = a;
      std::fwrite(&v, sizeof(v), 1, fo);
    }
    std::fclose(fo);
    return 0;
  }
  std::fprintf(stderr, "usage: dropin_test me|fwd|inv|mc|cavlc|fast ...\n");
  return 1;
}
