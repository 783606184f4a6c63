// codecutils_hip.hpp - C++ drop-in classes over the icu_hip C ABI.
//
// The classes keep the reference's class names, constructor arguments,
// two-stage Create(), method names and in-place buffer semantics, so codec
// code written against CodecUtils reads the same:
//   IMotionEstimator / MotionEstimatorH264ImplFull
//       include/CodecUtils/IMotionEstimator.h:51-96,
//       include/CodecUtils/MotionEstimatorH264ImplFull.h:72-221
//   IMotionVectorPredictor / H264MotionVectorPredictorImpl1
//       include/CodecUtils/IMotionVectorPredictor.h:51-83,
//       include/CodecUtils/H264MotionVectorPredictorImpl1.h:54-132
//   VectorStructList (SIMPLE2D)   include/CodecUtils/VectorStructList.h:56-110
//   IForwardTransform / FastForward4x4ITImpl1 / FastForward4x4On16x16ITImpl1
//       include/CodecUtils/IForwardTransform.h:56-126
//   IInverseTransform / FastInverse4x4ITImpl1 / FastInverse4x4On16x16ITImpl1
//       include/CodecUtils/IInverseTransform.h:56-115
//   YuvRawFileHandler   include/CodecUtils/YuvRawFileHandler.h:53-98
//   IMotionCompensator / MotionCompensatorH264ImplStd
//       include/CodecUtils/IMotionCompensator.h, MotionCompensatorH264ImplStd.h:55-110
//   MotionEstimatorH264ImplUMHS / ImplCross / ImplFHS
//       include/CodecUtils/MotionEstimatorH264ImplUMHS.h:78-126, ...Cross.h:78-190, ...FHS.h:78-214
//   IContextAwareRunLevelCodec / CAVLCH264Impl, BitStreamWriterMSB / BitStreamReaderMSB
//       include/CodecUtils/IContextAwareRunLevelCodec.h:49-103, CAVLCH264Impl.h:59-150,
//       BitStreamBaseMSB.h:62-87
// They live in namespace icu_codecutils; a codec that must keep its own
// CodecUtils interface types binds them through a thin adapter class that
// derives from its interface and forwards each call to these.
//
// Header only; link with libicu_hip.so.
#pragma once

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "icu_hip.h"

namespace icu_codecutils {

// ------------------------------------------------------------ MV container
class VectorStructList {
 public:
  static const int SIMPLE2D = 0;
  explicit VectorStructList(int type = SIMPLE2D) : _type(type) {}
  int SetLength(int length) {
    _v.assign(2 * (size_t)length, 0);
    return 1;
  }
  int GetLength(void) const { return (int)(_v.size() / 2); }
  int GetType(void) const { return _type; }
  int GetSimpleElement(int pos, int element) const { return _v[2 * (size_t)pos + element]; }
  void SetSimpleElement(int pos, int element, int value) { _v[2 * (size_t)pos + element] = (int16_t)value; }
  int16_t* Data(void) { return _v.data(); }
  // VectorStructList::Dump: "mv x,mv y" then one "x,y" line per vector (the
  // reference writes no title line)
  void Dump(char* filename, const char* title) {
    if (_type == SIMPLE2D) icu_mv_dump(filename, title, _v.data(), GetLength());
  }

 private:
  int _type;
  std::vector<int16_t> _v;
};

// ------------------------------------------------------------- predictors
// The reference interface, member for member
// (include/CodecUtils/IMotionVectorPredictor.h:51-83).
class IMotionVectorPredictor {
 public:
  virtual ~IMotionVectorPredictor() {}
  virtual int Create(void) = 0;
  virtual int Get16x16Prediction(void* pList, int blk, int* predX, int* predY) = 0;
  virtual int Get16x16Prediction(void* pList, int blk, int* predX, int* predY, int* distortion) { return 0; }
  virtual void Set16x16MotionVector(int blk, int mvX, int mvY) {}
  virtual void Set16x16MotionVector(int blk, int mvX, int mvY, int distortion) {}
};

// H.264 median predictor over the MBs of one picture (one slice).  The
// estimator decides the vectors on the device; afterwards it writes them
// back here with Set16x16MotionVector, as the reference estimator does, so
// Get16x16Prediction answers for later users.
class H264MotionVectorPredictorImpl1 : public IMotionVectorPredictor {
 public:
  H264MotionVectorPredictorImpl1(int mbCols, int mbRows)
      : _w(mbCols), _h(mbRows), _mv(2 * (size_t)mbCols * mbRows, 0), _d((size_t)mbCols * mbRows, 0) {}
  int Create(void) override { return 1; }
  void Set16x16MotionVector(int blk, int mvX, int mvY) override {
    _mv[2 * (size_t)blk] = mvX;
    _mv[2 * (size_t)blk + 1] = mvY;
  }
  void Set16x16MotionVector(int blk, int mvX, int mvY, int distortion) override {
    Set16x16MotionVector(blk, mvX, mvY);
    _d[blk] = distortion;
  }
  // The median with its distortion (MacroBlockH264.cpp:818-915): each
  // component takes the least distortion among the neighbours it equals;
  // the y median is compared with C's x component, as the reference does.
  int Get16x16Prediction(void*, int blk, int* predX, int* predY, int* distortion) override {
    const int col = blk % _w, row = blk / _w;
    int A[3] = {0, 0, 0}, B[3] = {0, 0, 0}, Cc[3] = {0, 0, 0};
    const bool hasB = row > 0, hasC = row > 0 && col + 1 < _w;
    if (col > 0) getd(blk - 1, A);
    if (hasB) getd(blk - _w, B);
    if (hasC) getd(blk - _w + 1, Cc);
    else if (col > 0 && row > 0) getd(blk - _w - 1, Cc);
    if (!hasB && !hasC)
      for (int i = 0; i < 3; i++) B[i] = Cc[i] = A[i];
    const int x = median(A[0], B[0], Cc[0]), y = median(A[1], B[1], Cc[1]);
    int xd = A[2], yd = A[2];
    if (x == B[0] && B[2] < xd) xd = B[2];
    if (x == Cc[0] && Cc[2] < xd) xd = Cc[2];
    if (y == B[1] && B[2] < yd) yd = B[2];
    if (y == Cc[0] && Cc[2] < yd) yd = Cc[2];
    *predX = x;
    *predY = y;
    *distortion = yd < xd ? yd : xd;
    return 1;
  }
  int Get16x16Prediction(void*, int blk, int* predX, int* predY) override {
    const int col = blk % _w, row = blk / _w;
    int A[2] = {0, 0}, B[2] = {0, 0}, Cc[2] = {0, 0};
    const bool hasB = row > 0, hasC = row > 0 && col + 1 < _w;
    if (col > 0) get(blk - 1, A);
    if (hasB) get(blk - _w, B);
    if (hasC) get(blk - _w + 1, Cc);
    else if (col > 0 && row > 0) get(blk - _w - 1, Cc);
    if (!hasB && !hasC) {
      B[0] = Cc[0] = A[0];
      B[1] = Cc[1] = A[1];
    }
    *predX = median(A[0], B[0], Cc[0]);
    *predY = median(A[1], B[1], Cc[1]);
    return 1;
  }

 private:
  void get(int blk, int* v) const {
    v[0] = _mv[2 * (size_t)blk];
    v[1] = _mv[2 * (size_t)blk + 1];
  }
  void getd(int blk, int* v) const {
    get(blk, v);
    v[2] = _d[blk];
  }
  static int median(int x, int y, int z) {
    int mn = y < z ? y : z, mx = y < z ? z : y;
    if (x < mn) mn = x;
    if (x > mx) mx = x;
    return x + y + z - mn - mx;
  }
  int _w, _h;
  std::vector<int> _mv, _d;
};

// Always predicts the zero vector.
class ZeroMotionVectorPredictor : public IMotionVectorPredictor {
 public:
  int Create(void) override { return 1; }
  int Get16x16Prediction(void*, int, int* predX, int* predY) override {
    *predX = 0;
    *predY = 0;
    return 1;
  }
};

// The predictors the device reproduces, by the predictor object's type: 1
// the H.264 median (H264MotionVectorPredictorImpl1), 0 the zero vector
// (ZeroMotionVectorPredictor), -1 anything else (the estimators' Create()
// then fails rather than predict differently from the object it was given).
inline int DevicePredictorKind(IMotionVectorPredictor* p) {
  if (dynamic_cast<H264MotionVectorPredictorImpl1*>(p)) return 1;
  if (dynamic_cast<ZeroMotionVectorPredictor*>(p)) return 0;
  return -1;
}

// ---------------------------------------------------------------- estimator
class IMotionEstimator {
 public:
  virtual ~IMotionEstimator() {}
  virtual int Create(void) = 0;
  virtual void Reset(void) = 0;
  virtual int Ready(void) = 0;
  virtual void SetMode(int mode) = 0;
  virtual int GetMode(void) = 0;
  virtual void* Estimate(const void* pSrc, const void* pRef, long* avgDistortion) = 0;
  virtual void* Estimate(long* avgDistortion) = 0;
  virtual void* Estimate(long* avgDistortion, void* param) { return Estimate(avgDistortion); }
};

// The full-search estimator.  pSrc / pRef are int16 planes of 8-bit
// samples, borrowed for the object's lifetime and re-read by every
// Estimate().  The returned list holds quarter-pel (x, y) per MB in raster
// order and stays valid until the next Estimate() or destruction.
// Beyond the reference's constructor, an optional icu_me_config selects the
// metric / window / filter / sub-pel depth (defaults: the shipped class).
class MotionEstimatorH264ImplFull : public IMotionEstimator {
 public:
  MotionEstimatorH264ImplFull(const void* pSrc, const void* pRef, int imgWidth, int imgHeight,
                              int motionRange, IMotionVectorPredictor* pMVPred,
                              void* pDistortionIncluded = nullptr, const icu_me_config* modes = nullptr)
      : _src((const int16_t*)pSrc), _ref((const int16_t*)pRef), _w(imgWidth), _h(imgHeight),
        _range(motionRange), _pred(pMVPred), _include((const bool*)pDistortionIncluded) {
    if (modes) {
      _cfg = *modes;
      _have_modes = true;
    }
  }
  ~MotionEstimatorH264ImplFull() override { Destroy(); }

  int Create(void) override {
    Destroy();
    _err.clear();
    const int kind = DevicePredictorKind(_pred);
    if (kind < 0) {
      // the reference dereferences the predictor (Full.cpp:368)
      _err = _pred ? "predictor type not reproduced on the device (H264MotionVectorPredictorImpl1 or "
                     "ZeroMotionVectorPredictor)"
                   : "null motion vector predictor";
      return 0;
    }
    if (!_have_modes) icu_me_config_reference(&_cfg, _w, _h, _range);
    _cfg.width = _w;
    _cfg.height = _h;
    _cfg.range = _range / 4;
    _cfg.predictor = kind;
    if (icu_me_create(&_cfg, &_h_me) != ICU_OK) return 0;
    _list.SetLength((_w / 16) * (_h / 16));
    _ready = 1;
    return 1;
  }
  void Reset(void) override {}
  int Ready(void) override { return _ready; }
  void SetMode(int mode) override { _mode = mode; }
  int GetMode(void) override { return _mode; }
  // As the reference, the pointer arguments are ignored: the planes given
  // at construction are estimated.
  void* Estimate(const void*, const void*, long* avgDistortion) override { return Estimate(avgDistortion); }
  void* Estimate(long* avgDistortion) override {
    if (!_ready) return nullptr;
    const int n = _list.GetLength();
    std::vector<uint8_t> inc;
    if (_include) {
      inc.resize(n);
      for (int i = 0; i < n; i++) inc[i] = _include[i] ? 1 : 0;
    }
    long avg = 0;
    if (icu_me_estimate(_h_me, _src, _ref, _include ? inc.data() : nullptr, _list.Data(), nullptr,
                        &avg) != ICU_OK)
      return nullptr;
    if (_pred)
      for (int i = 0; i < n; i++)
        _pred->Set16x16MotionVector(i, _list.GetSimpleElement(i, 0), _list.GetSimpleElement(i, 1));
    if (avgDistortion) *avgDistortion = avg;
    return (void*)&_list;
  }
  const char* LastError(void) const { return _err.empty() ? icu_last_error() : _err.c_str(); }

 private:
  void Destroy(void) {
    if (_h_me) icu_me_destroy(_h_me);
    _h_me = nullptr;
    _ready = 0;
  }
  const int16_t* _src;
  const int16_t* _ref;
  int _w, _h, _range;
  IMotionVectorPredictor* _pred;
  const bool* _include;
  icu_me_config _cfg{};
  bool _have_modes = false;
  icu_me* _h_me = nullptr;
  VectorStructList _list;
  int _ready = 0, _mode = 1;
  std::string _err;
};

// ---------------------------------------------------------- fast estimators
// MotionEstimatorH264ImplUMHS / ImplCross / ImplFHS
// (include/CodecUtils/MotionEstimatorH264ImplUMHS.h:78-126, ...Cross.h:78-190,
// ...FHS.h:78-214): the reference's constructors, Create(), SetMode (0
// quarter, 1 half, 2 full pel; default quarter) and Estimate().  UMHS and
// FHS take their starting point from the H.264 median predictor with
// distortion, which the device evaluates itself (pass an
// H264MotionVectorPredictorImpl1; the decided vectors and distortions are
// written back to it, as the reference does); Cross also takes
// ZeroMotionVectorPredictor.  Create() fails for any other predictor
// object.  UMHS's previous-frame macroblocks are given
// as their 16x16 vectors and intra flags instead of MacroBlockH264s.  A
// mode change takes effect at the next Estimate() with a new device handle
// (for FHS, a fresh half-sample cache).
class FastEstimatorBase : public IMotionEstimator {
 public:
  FastEstimatorBase(const void* pSrc, const void* pRef, int w, int h, int motionRange, IMotionVectorPredictor* pred,
                    void* include)
      : _src((const int16_t*)pSrc), _ref((const int16_t*)pRef), _w(w), _h(h), _range(motionRange), _pred(pred),
        _include((const bool*)include) {}
  int Create(void) override {
    Release();
    _list.SetLength((_w / 16) * (_h / 16));
    _dist.assign((size_t)(_w / 16) * (_h / 16), 0);
    const int kind = DevicePredictorKind(_pred);
    _ready = (_w % 16 == 0 && _h % 16 == 0 && _w > 0 && _h > 0 && (kind == 1 || (kind == 0 && ZeroPredictorOk())))
                 ? 1 : 0;
    return _ready;
  }
  void Reset(void) override {}
  int Ready(void) override { return _ready; }
  void SetMode(int mode) override { _mode = mode; }
  int GetMode(void) override { return _mode; }
  void* Estimate(const void*, const void*, long* avgDistortion) override { return Estimate(avgDistortion); }
  void* Estimate(long* avgDistortion) override {
    if (!_ready) return nullptr;
    if (_made_mode != _mode) {
      Release();
      if (Make(_mode) != ICU_OK) return nullptr;
      _made_mode = _mode;
    }
    const int n = _list.GetLength();
    std::vector<uint8_t> inc;
    if (_include) {
      inc.resize(n);
      for (int i = 0; i < n; i++) inc[i] = _include[i] ? 1 : 0;
    }
    long avg = 0;
    if (Run(_include ? inc.data() : nullptr, &avg) != ICU_OK) return nullptr;
    if (_pred)
      for (int i = 0; i < n; i++)
        _pred->Set16x16MotionVector(i, _list.GetSimpleElement(i, 0), _list.GetSimpleElement(i, 1), _dist[i]);
    if (avgDistortion) *avgDistortion = avg;
    return (void*)&_list;
  }
  const char* LastError(void) const { return icu_last_error(); }

 protected:
  virtual int Make(int mode) = 0;
  virtual int Run(const uint8_t* include, long* avg) = 0;
  virtual void Release(void) = 0;
  // whether the estimator also runs from the zero predictor
  virtual bool ZeroPredictorOk(void) const { return false; }
  const int16_t* _src;
  const int16_t* _ref;
  int _w, _h, _range;
  IMotionVectorPredictor* _pred;
  const bool* _include;
  VectorStructList _list;
  std::vector<int32_t> _dist;
  int _ready = 0, _mode = 0, _made_mode = -1;
};

class MotionEstimatorH264ImplUMHS : public FastEstimatorBase {
 public:
  MotionEstimatorH264ImplUMHS(const void* pSrc, const void* pRef, int imgWidth, int imgHeight, int motionRange,
                              IMotionVectorPredictor* pMVPred, void* pDistortionIncluded = nullptr,
                              const int16_t* pPrevMv = nullptr, const uint8_t* pPrevIntra = nullptr)
      : FastEstimatorBase(pSrc, pRef, imgWidth, imgHeight, motionRange, pMVPred, pDistortionIncluded),
        _prev(pPrevMv), _prev_intra(pPrevIntra) {}
  ~MotionEstimatorH264ImplUMHS() override { Release(); }

 protected:
  int Make(int mode) override { return icu_umhs_create(_w, _h, _range, mode, &_hd); }
  int Run(const uint8_t* include, long* avg) override {
    return icu_umhs_estimate(_hd, _src, _ref, _prev, _prev ? _prev_intra : nullptr, include, _list.Data(),
                             _dist.data(), avg);
  }
  void Release(void) override {
    if (_hd) icu_umhs_destroy(_hd);
    _hd = nullptr;
    _made_mode = -1;
  }

 private:
  const int16_t* _prev;
  const uint8_t* _prev_intra;
  icu_umhs* _hd = nullptr;
};

class MotionEstimatorH264ImplCross : public FastEstimatorBase {
 public:
  MotionEstimatorH264ImplCross(const void* pSrc, const void* pRef, int imgWidth, int imgHeight, int motionRange,
                               IMotionVectorPredictor* pMVPred, void* pDistortionIncluded = nullptr)
      : FastEstimatorBase(pSrc, pRef, imgWidth, imgHeight, motionRange, pMVPred, pDistortionIncluded) {}
  ~MotionEstimatorH264ImplCross() override { Release(); }

 protected:
  int Make(int mode) override { return icu_cross_create(_w, _h, _range, mode, DevicePredictorKind(_pred), &_hd); }
  bool ZeroPredictorOk(void) const override { return true; }
  int Run(const uint8_t* include, long* avg) override {
    return icu_cross_estimate(_hd, _src, _ref, include, _list.Data(), _dist.data(), avg);
  }
  void Release(void) override {
    if (_hd) icu_cross_destroy(_hd);
    _hd = nullptr;
    _made_mode = -1;
  }

 private:
  icu_cross* _hd = nullptr;
};

class MotionEstimatorH264ImplFHS : public FastEstimatorBase {
 public:
  MotionEstimatorH264ImplFHS(const void* pSrc, const void* pRef, int imgWidth, int imgHeight, int motionRange,
                             IMotionVectorPredictor* pMVPred, void* pDistortionIncluded = nullptr)
      : FastEstimatorBase(pSrc, pRef, imgWidth, imgHeight, motionRange, pMVPred, pDistortionIncluded) {}
  ~MotionEstimatorH264ImplFHS() override { Release(); }

 protected:
  int Make(int mode) override { return icu_fhs_create(_w, _h, _range, mode, &_hd); }
  int Run(const uint8_t* include, long* avg) override {
    return icu_fhs_estimate(_hd, _src, _ref, include, _list.Data(), _dist.data(), avg);
  }
  void Release(void) override {
    if (_hd) icu_fhs_destroy(_hd);
    _hd = nullptr;
    _made_mode = -1;
  }

 private:
  icu_fhs* _hd = nullptr;
};

// --------------------------------------------------------------- transforms
class IForwardTransform {
 public:
  static const int QUANT_ID = 0;
  static const int INTRA_FLAG_ID = 1;
  static const int TransformAndQuant = 0;
  static const int TransformOnly = 1;
  static const int QuantOnly = 2;
  virtual ~IForwardTransform() {}
  virtual void Transform(void* ptr) = 0;
  virtual void Transform(void* pIn, void* pCoeff) = 0;
  virtual void SetScale(void* pScale) = 0;
  virtual void* GetScale(void) = 0;
  virtual void SetMode(int mode) = 0;
  virtual int GetMode(void) = 0;
  virtual void SetParameter(int paramID, int paramVal) = 0;
  virtual int GetParameter(int paramID) = 0;
  virtual int QuantiseValue(short val, int pos, int qp) = 0;
};

class IInverseTransform {
 public:
  static const int QUANT_ID = 0;
  static const int TransformAndQuant = 0;
  static const int TransformOnly = 1;
  static const int QuantOnly = 2;
  virtual ~IInverseTransform() {}
  virtual void InverseTransform(void* ptr) = 0;
  virtual void InverseTransform(void* pCoeff, void* pOut) = 0;
  virtual void SetScale(void* pScale) = 0;
  virtual void* GetScale(void) = 0;
  virtual void SetMode(int mode) = 0;
  virtual int GetMode(void) = 0;
  virtual void SetParameter(int paramID, int paramVal) = 0;
  virtual int GetParameter(int paramID) = 0;
};

// Forward 4x4 IT + quant.  Layout 0 (one 4x4 block, 16 shorts per call) is
// FastForward4x4ITImpl1; layout 1 (a 16x16 raster MB) is
// FastForward4x4On16x16ITImpl1.  TransformBlocks(ptr, n) is the batched
// form: n calls' worth in one call.  Per-block calls run on the CPU with the
// device kernels' arithmetic (icu_fwd4x4's small-batch path); batches of
// more than 16384 blocks go to the device.
template <int LAYOUT>
class FastForward4x4ITHip : public IForwardTransform {
 public:
  void Transform(void* ptr) override { TransformBlocks(ptr, 1); }
  void Transform(void* pIn, void* pCoeff) override {
    std::memcpy(pCoeff, pIn, sizeof(short) * (LAYOUT ? 256 : 16));
    Transform(pCoeff);
  }
  void TransformBlocks(void* ptr, long calls) {
    const long nb = calls * (LAYOUT ? 16 : 1);
    icu_fwd4x4((const int16_t*)ptr, (int16_t*)ptr, nb, _q, _intra, _mode, LAYOUT);
  }
  void SetScale(void*) override {}
  void* GetScale(void) override { return nullptr; }
  void SetMode(int mode) override { _mode = mode; }
  int GetMode(void) override { return _mode; }
  void SetParameter(int id, int v) override {
    if (id == QUANT_ID) _q = v;
    else if (id == INTRA_FLAG_ID) _intra = v;
  }
  int GetParameter(int id) override { return id == QUANT_ID ? _q : _intra; }
  // Not implemented by the reference Impl1 either (FastForward4x4ITImpl1.h:118).
  int QuantiseValue(short, int, int) override { return 0; }

 private:
  int _mode = TransformAndQuant, _intra = 1, _q = 1;
};
typedef FastForward4x4ITHip<0> FastForward4x4ITImpl1;
// The reference's unrolled variant (src/CodecUtils/FastForward4x4ITImpl2.cpp:110-450)
// gives Impl1's output in every mode, so it is the same class here.
typedef FastForward4x4ITHip<0> FastForward4x4ITImpl2;
typedef FastForward4x4ITHip<1> FastForward4x4On16x16ITImpl1;

// Dequant + inverse 4x4 IT.  Layout 0 / rows first is FastInverse4x4ITImpl1;
// layout 1 / columns first is FastInverse4x4On16x16ITImpl1 (not bit-equal).
template <int LAYOUT>
class FastInverse4x4ITHip : public IInverseTransform {
 public:
  FastInverse4x4ITHip() {
    for (int i = 0; i < 16; i++) _w[i] = 16;
  }
  void InverseTransform(void* ptr) override { InverseTransformBlocks(ptr, 1); }
  void InverseTransform(void* pCoeff, void* pOut) override {
    std::memcpy(pOut, pCoeff, sizeof(short) * (LAYOUT ? 256 : 16));
    InverseTransform(pOut);
  }
  void InverseTransformBlocks(void* ptr, long calls) {
    const long nb = calls * (LAYOUT ? 16 : 1);
    icu_inv4x4((const int16_t*)ptr, (int16_t*)ptr, nb, _q, _mode, LAYOUT, LAYOUT, _w);
  }
  void SetScale(void* pScale) override { std::memcpy(_w, pScale, sizeof(_w)); }
  void* GetScale(void) override { return (void*)_w; }
  void SetMode(int mode) override { _mode = mode; }
  int GetMode(void) override { return _mode; }
  void SetParameter(int id, int v) override {
    if (id == QUANT_ID) _q = v;
  }
  int GetParameter(int) override { return _q; }

 private:
  int _mode = TransformAndQuant, _q = 1;
  int _w[16];
};
typedef FastInverse4x4ITHip<0> FastInverse4x4ITImpl1;
typedef FastInverse4x4ITHip<1> FastInverse4x4On16x16ITImpl1;

// ----------------------------------------------------- motion compensation
class IMotionCompensator {
 public:
  virtual ~IMotionCompensator() {}
  virtual int Create(void* ref, int imgWidth, int imgHeight, int macroBlkWidth, int macroBlkHeight) = 0;
  virtual void Destroy(void) = 0;
  virtual void Reset(void) = 0;
  virtual void Compensate(void* pMotionList) = 0;
};

// Compensates the YUV 4:2:0 reference buffer given to Create() in place, as
// the reference class: `range` only sized the reference's extended borders
// (here every vector is valid: samples past the picture repeat its edge).
// 16x16 macroblocks only.
class MotionCompensatorH264ImplStd : public IMotionCompensator {
 public:
  explicit MotionCompensatorH264ImplStd(int range) : _range(range) {}
  ~MotionCompensatorH264ImplStd() override { Destroy(); }
  int Create(void* ref, int imgWidth, int imgHeight, int macroBlkWidth, int macroBlkHeight) override {
    Destroy();
    if (!ref || macroBlkWidth != 16 || macroBlkHeight != 16) return 0;
    if (icu_mc_create(imgWidth, imgHeight, &_h) != ICU_OK) return 0;
    _ref = (int16_t*)ref;
    return 1;
  }
  void Destroy(void) override {
    if (_h) icu_mc_destroy(_h);
    _h = nullptr;
  }
  void Reset(void) override {}
  void Compensate(void* pMotionList) override {
    VectorStructList* l = (VectorStructList*)pMotionList;
    if (!_h || !l || l->GetLength() == 0 || l->GetType() != VectorStructList::SIMPLE2D) return;
    icu_mc_compensate(_h, _ref, l->Data());
  }
  const char* LastError(void) const { return icu_last_error(); }

 private:
  int _range;
  icu_mc* _h = nullptr;
  int16_t* _ref = nullptr;
};

// ---------------------------------------------------------------- raw YUV
// Reads a raw 4:2:0 file unit by unit (YUV4208P or YUV420NV12 in, YUV42016P
// or YUV4208P out) or collects units to write on Close(), as the reference
// handler; rw 0 reads, 1 writes.
class YuvRawFileHandler {
 public:
  static const int YUV4208P = ICU_YUV420_8P;
  static const int YUV420NV12 = ICU_YUV420_NV12;
  static const int YUV42016P = ICU_YUV420_16P;
  ~YuvRawFileHandler() { Close(); }
  void SetWidth(int width) { _w = width; }
  void SetHeight(int height) { _h = height; }
  void SetType(int type) { _in = type; }
  void SetType(int intype, int outtype) {
    _in = intype;
    _out = outtype;
  }
  int Open(const char* filename, int rw) {
    Close();
    return icu_yuv_file_open(filename, rw, _w, _h, _in, _out, &_f) == ICU_OK ? 1 : 0;
  }
  void Close(void) {
    if (_f) icu_yuv_file_close(_f);
    _f = nullptr;
    _unit = nullptr;
    _len = 0;
  }
  void* GetNextUnit(int* unitLen) {
    *unitLen = 0;
    if (!_f || icu_yuv_file_next_unit(_f, (const void**)&_unit, unitLen) != ICU_OK) return nullptr;
    _len = *unitLen;
    return _unit;
  }
  void* GetCurrUnit(int* unitLen) {
    *unitLen = _w * _h * 3 / 2;
    return _unit;
  }
  int PutNextUnit(void* unit, int unitLen) {
    return _f && icu_yuv_file_put_unit(_f, unit, unitLen) == ICU_OK ? 1 : 0;
  }

 private:
  int _w = 352, _h = 288, _in = YUV4208P, _out = YUV42016P;
  icu_yuv_file* _f = nullptr;
  void* _unit = nullptr;
  int _len = 0;
};

// ------------------------------------------------------- MSB-first streams
// A cursor over a caller's byte array, bits taken from the MSB of each byte
// down.  The cursor is kept linear; GetStreamBitPos() and Seek() exchange
// it in the reference base class's own encoding, (byte << 3) + bit with bit
// counting 7 (MSB) down to 0, so that Seek(GetStreamBitPos()) returns to the
// same place as with the reference (include/CodecUtils/BitStreamBaseMSB.h:62-87).
class BitStreamBaseMSB {
 public:
  virtual ~BitStreamBaseMSB() {}
  virtual void SetStream(void* stream, int bitSize) {
    _s = (uint8_t*)stream;
    _size = bitSize;
    _pos = 0;
  }
  virtual void* GetStream(void) { return _s; }
  virtual void Reset(void) { _pos = 0; }
  virtual int Seek(int streamBitPos) {
    if (streamBitPos >= _size) return 0;
    _pos = (long)(streamBitPos / 8) * 8 + 7 - streamBitPos % 8;
    return 1;
  }
  virtual int GetStreamBitPos(void) { return (int)((_pos >> 3) << 3) + 7 - (int)(_pos & 7); }
  virtual int GetStreamBytePos(void) { return (int)(_pos >> 3); }
  virtual void SetStreamBitSize(int bitSize) { _size = bitSize; }
  virtual int GetStreamBitSize(void) { return _size; }
  virtual int GetStreamBitsRemaining(void) { return _size - (int)_pos; }
  // the linear position, for the entropy coder
  long& Cursor(void) { return _pos; }
  uint8_t* Bytes(void) { return _s; }

 protected:
  uint8_t* _s = nullptr;
  int _size = 0;
  long _pos = 0;
};

// include/CodecUtils/BitStreamWriterMSB.h (src/CodecUtils/BitStreamWriterMSB.cpp:73-180):
// each written bit is set or cleared, the rest of the byte kept.
class BitStreamWriterMSB : public BitStreamBaseMSB {
 public:
  void Write(int val) { Write(1, val); }
  void Write(int numBits, int val) {
    for (int i = numBits - 1; i >= 0; i--, _pos++) put(_pos, (val >> i) & 1);
  }
  void Write(int numBits, int valH, int valL) {
    if (numBits > 32) {
      Write(numBits - 32, valH);
      Write(32, valL);
    } else {
      Write(numBits, valL);
    }
  }
  void Poke(int bitLoc, int numBits, int val) {
    // bitLoc in the same encoding as Seek()
    long p = (long)(bitLoc / 8) * 8 + 7 - bitLoc % 8;
    for (int i = numBits - 1; i >= 0; i--, p++) put(p, (val >> i) & 1);
  }

 private:
  void put(long p, int b) {
    const uint8_t m = (uint8_t)(0x80u >> (p & 7));
    if (b) _s[p >> 3] |= m;
    else _s[p >> 3] &= (uint8_t)~m;
  }
};

// include/CodecUtils/BitStreamReaderMSB.h (src/CodecUtils/BitStreamReaderMSB.cpp:71-160).
class BitStreamReaderMSB : public BitStreamBaseMSB {
 public:
  int Read(void) { return Read(1); }
  int Read(int numBits) {
    int v = 0;
    for (int i = 0; i < numBits; i++, _pos++) v = (v << 1) | ((_s[_pos >> 3] >> (7 - (_pos & 7))) & 1);
    return v;
  }
  int Peek(int bitLoc, int numBits) {
    long p = (long)(bitLoc / 8) * 8 + 7 - bitLoc % 8;
    int v = 0;
    for (int i = 0; i < numBits; i++, p++) v = (v << 1) | ((_s[p >> 3] >> (7 - (p & 7))) & 1);
    return v;
  }
};

// ------------------------------------------------------------------ CAVLC
// IContextAwareRunLevelCodec / CAVLCH264Impl
// (include/CodecUtils/IContextAwareRunLevelCodec.h:49-103,
// include/CodecUtils/CAVLCH264Impl.h:59-150): Encode(short block[], writer)
// and Decode(reader, short block[]) return the bits coded or
// VLC_SYMBOL_NOT_RECOGNISED / STREAM_ACCESS_DENIED; the stream arguments are
// this header's BitStreamWriterMSB / BitStreamReaderMSB.  Mode8x8 is not
// coded (H.264 codes an 8x8 transform's levels as four 4x4 blocks) and
// returns VLC_SYMBOL_NOT_RECOGNISED.
class IContextAwareRunLevelCodec {
 public:
  virtual ~IContextAwareRunLevelCodec() {}
  virtual int Encode(void* in, void* stream) = 0;
  virtual int Decode(void* stream, void* out) = 0;
  virtual void SetMode(int mode) = 0;
  virtual int GetMode(void) = 0;
  virtual void SetParameter(int paramID, int paramVal) = 0;
  virtual int GetParameter(int paramID) = 0;
  static const int Mode2x2 = 0;
  static const int Mode4x4 = 1;
  static const int Mode8x8 = 2;
  static const int NUM_TOT_NEIGHBOR_COEFF_ID = 0;
  static const int DC_SKIP_FLAG_ID = 1;
  static const int NUM_TOT_COEFF_ID = 2;
};

class CAVLCH264Impl : public IContextAwareRunLevelCodec {
 public:
  static const int VLC_SYMBOL_NOT_RECOGNISED = -1;
  static const int STREAM_ACCESS_DENIED = -2;
  // Mode8x8: 64 levels in the 8x8 zig-zag scan, coded with the 4x4 tables
  // as the reference does (icu_cavlc_encode says which blocks code).  On an
  // error the writer keeps the symbols written before it, as the
  // reference's does.
  int Encode(void* in, void* stream) override {
    BitStreamWriterMSB* w = (BitStreamWriterMSB*)stream;
    int nbits = 0;
    const int rc = icu_cavlc_encode((const int16_t*)in, _mode, _nc, _dc, w->Bytes(), w->GetStreamBitSize(),
                                    &w->Cursor(), &nbits);
    if (rc != ICU_OK) return rc == ICU_E_RANGE ? STREAM_ACCESS_DENIED : VLC_SYMBOL_NOT_RECOGNISED;
    _num = count((const int16_t*)in);
    return nbits;
  }
  int Decode(void* stream, void* out) override {
    BitStreamReaderMSB* r = (BitStreamReaderMSB*)stream;
    int nbits = 0;
    const int rc = icu_cavlc_decode(r->Bytes(), r->GetStreamBitSize(), &r->Cursor(), _mode, _nc, _dc,
                                    (int16_t*)out, &nbits);
    if (rc != ICU_OK) return rc == ICU_E_RANGE ? STREAM_ACCESS_DENIED : VLC_SYMBOL_NOT_RECOGNISED;
    _num = count((const int16_t*)out);
    return nbits;
  }
  // unknown modes select 4x4, as the reference
  void SetMode(int mode) override { _mode = (mode == Mode2x2 || mode == Mode8x8) ? mode : Mode4x4; }
  int GetMode(void) override { return _mode; }
  void SetParameter(int paramID, int paramVal) override {
    if (paramID == NUM_TOT_NEIGHBOR_COEFF_ID) _nc = paramVal;
    else if (paramID == DC_SKIP_FLAG_ID) _dc = paramVal;
  }
  int GetParameter(int paramID) override {
    return paramID == NUM_TOT_NEIGHBOR_COEFF_ID ? _nc : paramID == DC_SKIP_FLAG_ID ? _dc
                                                    : paramID == NUM_TOT_COEFF_ID ? _num : -1;
  }

 private:
  // levels coded in the block: every position but a skipped DC
  int count(const int16_t* b) const {
    int n = 0;
    for (int i = _dc ? 1 : 0; i < (_mode == Mode2x2 ? 4 : (_mode == Mode8x8 ? 64 : 16)); i++) n += b[i] != 0;
    return n;
  }
  int _mode = Mode4x4, _nc = 0, _dc = 0, _num = 0;
};

}  // namespace icu_codecutils
