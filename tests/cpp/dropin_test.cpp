// Drives the C++ drop-in classes the way codec code drives the reference
// classes (construct, Create, Estimate, read the VectorStructList; transform
// objects with SetParameter / SetMode / Transform in place).  Inputs and
// outputs are raw files; tests/test_gpu_dropin.py compares them.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
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
    const int size = mode == IContextAwareRunLevelCodec::Mode2x2 ? 4 : (mode == IContextAwareRunLevelCodec::Mode8x8 ? 64 : 16);
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
  if (cmd == "cavlc_each" && argc >= 8) {
    // cavlc_each mode dc n blocks.raw nc.raw out.raw: every block into a
    // writer of its own (256 zeroed bytes), so a failed Encode cannot shift
    // the next block; out.raw holds per block Encode's return, the bits the
    // writer holds afterwards, NUM_TOT_COEFF, the 256 stream bytes, then (a
    // block that encoded) Decode's return and the decoded levels.
    const int mode = std::atoi(argv[2]), dc = std::atoi(argv[3]), n = std::atoi(argv[4]);
    const int size = mode == IContextAwareRunLevelCodec::Mode2x2 ? 4 : (mode == IContextAwareRunLevelCodec::Mode8x8 ? 64 : 16);
    auto blocks = load<short>(argv[5], (size_t)n * size);
    auto nc = load<signed char>(argv[6], (size_t)n);
    CAVLCH264Impl codec;
    codec.SetMode(mode);
    codec.SetParameter(IContextAwareRunLevelCodec::DC_SKIP_FLAG_ID, dc);
    std::vector<int> out;
    for (int i = 0; i < n; i++) {
      std::vector<unsigned char> stream(256, 0);
      BitStreamWriterMSB w;
      w.SetStream(stream.data(), 256 * 8);
      codec.SetParameter(IContextAwareRunLevelCodec::NUM_TOT_NEIGHBOR_COEFF_ID, nc[i]);
      const int e = codec.Encode(&blocks[(size_t)i * size], &w);
      out.push_back(e);
      out.push_back(256 * 8 - w.GetStreamBitsRemaining());
      out.push_back(codec.GetParameter(IContextAwareRunLevelCodec::NUM_TOT_COEFF_ID));
      for (unsigned char c : stream) out.push_back(c);
      std::vector<short> dec(size, 0);
      int d = 0;
      if (e > 0) {
        BitStreamReaderMSB r;
        r.SetStream(stream.data(), 256 * 8);
        d = codec.Decode(&r, dec.data());
      }
      out.push_back(d);
      for (int k = 0; k < size; k++) out.push_back(dec[k]);
    }
    save(argv[7], out.data(), out.size());
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
  if (cmd == "api") {
    // The interface points a codec relies on beyond the fixtures: which
    // predictors Create() accepts, the Estimate(avg, param) overload, the
    // Impl2 transform name.
    const int W = 64, H = 48;
    std::vector<short> cur((size_t)W * H), ref((size_t)W * H);
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) {
        ref[(size_t)y * W + x] = (short)((x * 7 + y * 13 + ((x * y) % 11)) & 255);
        cur[(size_t)y * W + x] = ref[(size_t)y * W + ((x + 3) % W)];
      }
    struct OtherPredictor : IMotionVectorPredictor {
      int Create(void) override { return 1; }
      int Get16x16Prediction(void*, int, int* x, int* y) override {
        *x = 4;
        *y = 4;
        return 1;
      }
    } other;
    MotionEstimatorH264ImplFull none(cur.data(), ref.data(), W, H, 32, nullptr);
    if (none.Create()) return 10;
    MotionEstimatorH264ImplFull odd(cur.data(), ref.data(), W, H, 32, &other);
    if (odd.Create() || std::string(odd.LastError()).find("predictor") == std::string::npos) return 11;
    ZeroMotionVectorPredictor zero;
    MotionEstimatorH264ImplFull me(cur.data(), ref.data(), W, H, 32, &zero);
    if (!me.Create()) return 12;
    IMotionEstimator* ime = &me;
    long a1 = -1, a2 = -1;
    VectorStructList* l1 = (VectorStructList*)ime->Estimate(&a1);
    std::vector<int> v1;
    for (int i = 0; l1 && i < l1->GetLength(); i++) v1.push_back(l1->GetSimpleElement(i, 0) * 65536 + l1->GetSimpleElement(i, 1));
    VectorStructList* l2 = (VectorStructList*)ime->Estimate(&a2, (void*)&other);
    if (!l1 || !l2 || a1 != a2) return 13;
    for (int i = 0; i < l2->GetLength(); i++)
      if (v1[i] != l2->GetSimpleElement(i, 0) * 65536 + l2->GetSimpleElement(i, 1)) return 14;
    // UMHS / FHS start from the median with distortion; Cross also from zero
    MotionEstimatorH264ImplUMHS um(cur.data(), ref.data(), W, H, 32, &zero);
    MotionEstimatorH264ImplFHS fh(cur.data(), ref.data(), W, H, 32, &other);
    MotionEstimatorH264ImplCross cr(cur.data(), ref.data(), W, H, 32, &zero);
    MotionEstimatorH264ImplCross cr2(cur.data(), ref.data(), W, H, 32, &other);
    if (um.Create() || fh.Create() || !cr.Create() || cr2.Create()) return 15;
    // FastForward4x4ITImpl2 is Impl1's transform in every mode
    for (int mode = 0; mode < 3; mode++)
      for (int qp = 0; qp < 52; qp += 3) {
        short b1[16], b2[16];
        for (int k = 0; k < 16; k++) b1[k] = b2[k] = (short)(((k * 37 + qp * 11 + mode * 5) % 511) - 255);
        FastForward4x4ITImpl1 t1;
        FastForward4x4ITImpl2 t2;
        IForwardTransform* ft = &t2;
        t1.SetParameter(IForwardTransform::QUANT_ID, qp);
        ft->SetParameter(IForwardTransform::QUANT_ID, qp);
        t1.SetMode(mode);
        ft->SetMode(mode);
        t1.Transform(b1);
        ft->Transform(b2);
        if (std::memcmp(b1, b2, sizeof(b1))) return 16;
      }
    std::printf("ok\n");
    return 0;
  }
  if (cmd == "txtime" && argc >= 3) {
    // txtime n: n single-block Transform() calls, then n InverseTransform()
    // calls, as the reference's BlockH264 makes them (one per 4x4 block),
    // checked against the batched call; prints ns per call.
    const long n = std::atol(argv[2]);
    std::vector<short> x((size_t)n * 16), y;
    for (size_t i = 0; i < x.size(); i++) x[i] = (short)(((i * 2654435761u) >> 7) % 511) - 255;
    y = x;
    std::vector<short> z = x;
    FastForward4x4ITImpl1 f;
    FastInverse4x4ITImpl1 inv;
    f.SetParameter(IForwardTransform::QUANT_ID, 28);
    f.SetParameter(IForwardTransform::INTRA_FLAG_ID, 0);
    inv.SetParameter(IInverseTransform::QUANT_ID, 28);
    const auto t0 = std::chrono::steady_clock::now();
    for (long i = 0; i < n; i++) f.Transform(y.data() + 16 * i);
    const auto t1 = std::chrono::steady_clock::now();
    for (long i = 0; i < n; i++) inv.InverseTransform(y.data() + 16 * i);
    const auto t2 = std::chrono::steady_clock::now();
    f.TransformBlocks(z.data(), n);
    inv.InverseTransformBlocks(z.data(), n);
    if (y != z) return 20;
    const double fns = std::chrono::duration<double, std::nano>(t1 - t0).count() / n;
    const double ins = std::chrono::duration<double, std::nano>(t2 - t1).count() / n;
    std::printf("fwd_ns_per_call %.1f inv_ns_per_call %.1f\n", fns, ins);
    return 0;
  }
  if (cmd == "threads" && argc >= 7) {
    // threads umhs|fhs W H nthreads nframes: one estimator object per
    // thread (each its own stream of frames), as a codec coding several
    // streams on several host threads drives the reference classes; first
    // all streams one after another, then all threads at once.  Results
    // must be the same; prints the wall time per frame of both.
    const std::string kind = argv[2];
    const int W = std::atoi(argv[3]), H = std::atoi(argv[4]), nt = std::atoi(argv[5]), nf = std::atoi(argv[6]);
    const size_t np = (size_t)W * H;
    const int nmb = (W / 16) * (H / 16);
    // frame f of stream t: a texture panned by (t + f, 2 f) pels
    auto frame = [&](int t, int f, std::vector<short>& out) {
      for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
          const int u = x + t + f, v = y + 2 * f;
          out[(size_t)y * W + x] = (short)((u * 7 + v * 13 + ((u * v) % 23) + ((u / 9 + v / 7) % 5) * 31) & 255);
        }
    };
    auto run = [&](int t, std::vector<short>* out) {
      std::vector<short> cur(np), ref(np);
      H264MotionVectorPredictorImpl1 med(W / 16, H / 16);
      IMotionEstimator* me;
      if (kind == "umhs") me = new MotionEstimatorH264ImplUMHS(cur.data(), ref.data(), W, H, 128, &med);
      else me = new MotionEstimatorH264ImplFHS(cur.data(), ref.data(), W, H, 128, &med);
      if (!me->Create()) {
        delete me;
        return false;
      }
      me->SetMode(0);
      bool ok = true;
      for (int f = 0; f < nf && ok; f++) {
        frame(t, f, ref);
        frame(t, f + 1, cur);
        long avg = 0;
        VectorStructList* l = (VectorStructList*)me->Estimate(&avg);
        if (!l) {
          ok = false;
          break;
        }
        for (int i = 0; i < nmb; i++) {
          out->push_back((short)l->GetSimpleElement(i, 0));
          out->push_back((short)l->GetSimpleElement(i, 1));
        }
      }
      delete me;
      return ok;
    };
    std::vector<std::vector<short>> serial(nt), par(nt);
    run(0, &serial[0]);  // warm-up: device initialisation, first launches
    serial[0].clear();
    const auto t0 = std::chrono::steady_clock::now();
    for (int t = 0; t < nt; t++)
      if (!run(t, &serial[t])) return 30;
    const auto t1 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    std::vector<int> oks(nt, 0);
    for (int t = 0; t < nt; t++) th.emplace_back([&, t] { oks[t] = run(t, &par[t]) ? 1 : 0; });
    for (auto& x : th) x.join();
    const auto t2 = std::chrono::steady_clock::now();
    for (int t = 0; t < nt; t++) {
      if (!oks[t]) return 31;
      if (par[t] != serial[t]) return 32;
    }
    const double frames = (double)nt * nf;
    std::printf("serial_ms_per_frame %.3f threaded_ms_per_frame %.3f\n",
                std::chrono::duration<double, std::milli>(t1 - t0).count() / frames,
                std::chrono::duration<double, std::milli>(t2 - t1).count() / frames);
    return 0;
  }
  std::fprintf(stderr, "usage: dropin_test me|fwd|inv|mc|cavlc|fast|api|txtime|threads ...\n");
  return 1;
}
