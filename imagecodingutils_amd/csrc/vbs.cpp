// C ABI of the multi-reference variable-block-size estimator and its
// residual coder (icu_vbs_*).  No reference counterpart: the reference's
// estimators search 16x16 blocks of one reference frame.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/icu_hip.h"
#include "icu_device.h"
#include "icu_internal.h"

using namespace icu;

struct icu_vbs {
  icu_vbs_config cfg{};
  VbsParams plan{};
  uint8_t* refpad_base[kVbsMaxRefs] = {};
  uint8_t* refpad[kVbsMaxRefs] = {};
  int ref_stride = 0;
  int pad_wgs = 0;  // workgroups padding one plane
  hipStream_t own = nullptr;
  ~icu_vbs() {
    for (uint8_t* p : refpad_base) hipFree(p);
    if (own) hipStreamDestroy(own);
  }
};

namespace {

hipStream_t pick(void* stream, icu_vbs* h) { return stream ? (hipStream_t)stream : h->own; }

#define VBS_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return icu_fail(ICU_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

extern "C" {

int icu_vbs_create(const icu_vbs_config* cfg, icu_vbs** out) {
  if (!cfg || !out) return icu_fail(ICU_E_ARG, "null argument");
  const icu_vbs_config& c = *cfg;
  if (c.width < 16 || c.height < 16 || c.width % 16 || c.height % 16)
    return icu_fail(ICU_E_ARG, "width and height must be positive multiples of 16");
  if (c.nrefs < 1 || c.nrefs > kVbsMaxRefs) return icu_fail(ICU_E_ARG, "nrefs must be 1..4");
  if (c.range < 1 || c.range > 32) return icu_fail(ICU_E_ARG, "range must be 1..32");
  // keys hold cost << 15 | index: cost (SAD <= 65280 plus the rate) below
  // 2^17, index (nrefs (2R+1)^2) below 2^15
  if (c.lambda < 0 || 65280L + 2L * c.range * c.lambda >= (1L << 17))
    return icu_fail(ICU_E_ARG, "lambda * 2 * range must stay below 65792");
  icu_vbs* h = new icu_vbs();
  h->cfg = c;
  vbs_plan(c.range, c.nrefs, &h->plan);
  h->plan.W = c.width;
  h->plan.H = c.height;
  h->plan.mbw = c.width / 16;
  h->plan.nmb = (c.width / 16) * (c.height / 16);
  h->plan.lambda = c.lambda;
  if (vbs_lds_bytes(h->plan) > 64 * 1024) {
    delete h;
    return icu_fail(ICU_E_ARG, "search window does not fit the workgroup's LDS");
  }
  h->ref_stride = ((c.width + 2 * kPad) + 255) & ~255;
  const size_t bytes = (size_t)h->ref_stride * (c.height + 2 * kPad);
  for (int i = 0; i < c.nrefs; i++) {
    if (hipMalloc(&h->refpad_base[i], bytes) != hipSuccess) {
      delete h;
      return icu_fail(ICU_E_DEVICE, "device allocation failed");
    }
    h->refpad[i] = h->refpad_base[i] + (size_t)kPad * h->ref_stride + kPad;
  }
  if (hipStreamCreate(&h->own) != hipSuccess) {
    delete h;
    return icu_fail(ICU_E_DEVICE, "hipStreamCreate failed");
  }
  h->pad_wgs = pad_plane_workgroups(c.width, c.height);
  h->plan.ref_stride = h->ref_stride;
  for (int i = 0; i < kVbsMaxRefs; i++) h->plan.ref[i] = h->refpad[i < c.nrefs ? i : 0];
  *out = h;
  return ICU_OK;
}

int icu_vbs_destroy(icu_vbs* h) {
  delete h;
  return ICU_OK;
}

int icu_vbs_estimate_device(icu_vbs* h, const uint8_t* d_cur, const uint8_t* const* d_refs,
                            uint8_t* d_mode, int16_t* d_mv8, uint8_t* d_ref8, int32_t* d_cost,
                            void* stream) {
  if (!h || !d_cur || !d_refs || !d_mode || !d_mv8 || !d_ref8 || !d_cost)
    return icu_fail(ICU_E_ARG, "null argument");
  if ((uintptr_t)d_cur & 3) return icu_fail(ICU_E_ARG, "d_cur must be 4-byte aligned");
  PadSet ps{};
  ps.n = h->cfg.nrefs;
  for (int i = 0; i < h->cfg.nrefs; i++) {
    if (!d_refs[i]) return icu_fail(ICU_E_ARG, "null reference plane");
    ps.d[i] = PadDesc{d_refs[i], h->refpad[i], h->cfg.width, h->cfg.height, h->cfg.width, h->ref_stride,
                      i * h->pad_wgs};
  }
  const hipStream_t st = pick(stream, h);
  launch_pad_set(ps, h->cfg.nrefs * h->pad_wgs, st);
  VbsParams p = h->plan;
  p.cur = d_cur;
  p.mode = d_mode;
  p.mv8 = d_mv8;
  p.ref8 = d_ref8;
  p.cost = d_cost;
  launch_vbs_search(p, st);
  VBS_TRY(hipGetLastError());
  return ICU_OK;
}

int icu_vbs_code_device(icu_vbs* h, const uint8_t* d_cur, const int16_t* d_mv8, const uint8_t* d_ref8,
                        int nqp, const int* qps, int intra, int cols_first, int16_t* const* d_levels,
                        uint8_t* const* d_recon, int16_t* d_pred, void* stream) {
  if (!h || !d_cur || !d_mv8 || !d_ref8 || !qps || !d_levels || !d_recon)
    return icu_fail(ICU_E_ARG, "null argument");
  if (nqp < 1 || nqp > kMaxQpSet) return icu_fail(ICU_E_ARG, "nqp must be 1..20");
  if (((uintptr_t)d_cur & 3) || ((uintptr_t)d_pred & 7))
    return icu_fail(ICU_E_ARG, "d_cur must be 4-byte, d_pred 8-byte aligned");
  QpSet qs{};
  qs.n = nqp;
  for (int i = 0; i < nqp; i++) {
    if (qps[i] < 0 || qps[i] > 51) return icu_fail(ICU_E_ARG, "qp must be 0..51");
    if (!d_levels[i] || !d_recon[i]) return icu_fail(ICU_E_ARG, "null output");
    if (((uintptr_t)d_recon[i] & 3) || ((uintptr_t)d_levels[i] & 7))
      return icu_fail(ICU_E_ARG, "d_recon must be 4-byte, d_levels 8-byte aligned");
    qs.qp[i] = qps[i];
    qs.levels[i] = d_levels[i];
    qs.recon[i] = d_recon[i];
  }
  VbsRefs refs{};
  for (int i = 0; i < kVbsMaxRefs; i++) refs.ref[i] = h->plan.ref[i];
  refs.stride = h->ref_stride;
  launch_code_vbs(d_cur, refs, h->cfg.width, h->cfg.height, d_mv8, d_ref8, qs, intra ? 1 : 0,
                  cols_first ? 1 : 0, d_pred, pick(stream, h));
  VBS_TRY(hipGetLastError());
  return ICU_OK;
}

int icu_vbs_get_plan(icu_vbs* h, int* tile_rows, int* lds_bytes) {
  if (!h) return icu_fail(ICU_E_ARG, "null handle");
  if (tile_rows) *tile_rows = vbs_tile_rows(h->plan);
  if (lds_bytes) *lds_bytes = vbs_lds_bytes(h->plan);
  return ICU_OK;
}

}  // extern "C"
