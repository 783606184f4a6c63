// C ABI of the whole-frame YUV 4:2:0 motion compensator (icu_mc_*), the
// device form of MotionCompensatorH264ImplStd
// (src/CodecUtils/MotionCompensatorH264ImplStd.cpp:76-473).
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/icu_hip.h"
#include "icu_device.h"
#include "icu_internal.h"

using namespace icu;

struct icu_mc {
  int W = 0, H = 0;
  uint8_t* pad_base[3] = {};  // padded Y, U, V
  uint8_t* pad[3] = {};       // pixel (0,0) of each
  int stride[3] = {};
  int pad_wgs[3] = {};
  uint8_t* d_plane[3] = {};   // staging for the host entry point
  int16_t* d_mv = nullptr;
  std::vector<uint8_t> h8;
  hipStream_t own = nullptr;
  ~icu_mc() {
    for (int i = 0; i < 3; i++) {
      hipFree(pad_base[i]);
      hipFree(d_plane[i]);
    }
    hipFree(d_mv);
    if (own) hipStreamDestroy(own);
  }
};

namespace {

#define MC_TRY(expr)                                                                                    \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess) return icu_fail(ICU_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

extern "C" {

int icu_mc_create(int width, int height, icu_mc** out) {
  if (!out) return icu_fail(ICU_E_ARG, "null argument");
  if (width < 16 || height < 16 || width % 16 || height % 16)
    return icu_fail(ICU_E_ARG, "width and height must be positive multiples of 16");
  icu_mc* h = new icu_mc();
  h->W = width;
  h->H = height;
  for (int i = 0; i < 3; i++) {
    const int w = i ? width / 2 : width, hh = i ? height / 2 : height;
    h->stride[i] = ((w + 2 * kPad) + 255) & ~255;
    h->pad_wgs[i] = pad_plane_workgroups(w, hh);
    if (hipMalloc(&h->pad_base[i], (size_t)h->stride[i] * (hh + 2 * kPad)) != hipSuccess ||
        hipMalloc(&h->d_plane[i], (size_t)w * hh) != hipSuccess) {
      delete h;
      return icu_fail(ICU_E_DEVICE, "device allocation failed");
    }
    h->pad[i] = h->pad_base[i] + (size_t)kPad * h->stride[i] + kPad;
  }
  if (hipMalloc(&h->d_mv, (size_t)(width / 16) * (height / 16) * 4) != hipSuccess ||
      hipStreamCreate(&h->own) != hipSuccess) {
    delete h;
    return icu_fail(ICU_E_DEVICE, "device allocation failed");
  }
  *out = h;
  return ICU_OK;
}

int icu_mc_destroy(icu_mc* h) {
  delete h;
  return ICU_OK;
}

int icu_mc_compensate_device(icu_mc* h, const uint8_t* d_ref_y, const uint8_t* d_ref_u,
                             const uint8_t* d_ref_v, const int16_t* d_mv, uint8_t* d_y, uint8_t* d_u,
                             uint8_t* d_v, void* stream) {
  if (!h || !d_ref_y || !d_ref_u || !d_ref_v || !d_mv || !d_y || !d_u || !d_v)
    return icu_fail(ICU_E_ARG, "null argument");
  if (((uintptr_t)d_y | (uintptr_t)d_u | (uintptr_t)d_v) & 3)
    return icu_fail(ICU_E_ARG, "output planes must be 4-byte aligned");
  const hipStream_t st = stream ? (hipStream_t)stream : h->own;
  PadSet ps{};
  ps.n = 3;
  const uint8_t* src[3] = {d_ref_y, d_ref_u, d_ref_v};
  int base = 0;
  for (int i = 0; i < 3; i++) {
    const int w = i ? h->W / 2 : h->W, hh = i ? h->H / 2 : h->H;
    ps.d[i] = PadDesc{src[i], h->pad[i], w, hh, w, h->stride[i], base};
    base += h->pad_wgs[i];
  }
  launch_pad_set(ps, base, st);
  launch_mc_yuv(h->pad[0], h->stride[0], h->pad[1], h->pad[2], h->stride[1], h->W, h->H, d_mv, d_y, d_u, d_v,
                st);
  MC_TRY(hipGetLastError());
  return ICU_OK;
}

int icu_mc_compensate(icu_mc* h, int16_t* yuv, const int16_t* mv) {
  if (!h || !yuv || !mv) return icu_fail(ICU_E_ARG, "null argument");
  const size_t ny = (size_t)h->W * h->H, nc = ny / 4, n = ny + 2 * nc;
  h->h8.resize(n);
  for (size_t i = 0; i < n; i++) {
    if (yuv[i] < 0 || yuv[i] > 255) return icu_fail(ICU_E_RANGE, "input sample outside 0..255");
    h->h8[i] = (uint8_t)yuv[i];
  }
  const size_t off[3] = {0, ny, ny + nc}, len[3] = {ny, nc, nc};
  for (int i = 0; i < 3; i++)
    MC_TRY(hipMemcpyAsync(h->d_plane[i], h->h8.data() + off[i], len[i], hipMemcpyHostToDevice, h->own));
  MC_TRY(hipMemcpyAsync(h->d_mv, mv, (size_t)(h->W / 16) * (h->H / 16) * 4, hipMemcpyHostToDevice, h->own));
  // in place, as the reference class compensates its reference buffer
  int rc = icu_mc_compensate_device(h, h->d_plane[0], h->d_plane[1], h->d_plane[2], h->d_mv, h->d_plane[0],
                                    h->d_plane[1], h->d_plane[2], h->own);
  if (rc) return rc;
  for (int i = 0; i < 3; i++)
    MC_TRY(hipMemcpyAsync(h->h8.data() + off[i], h->d_plane[i], len[i], hipMemcpyDeviceToHost, h->own));
  MC_TRY(hipStreamSynchronize(h->own));
  for (size_t i = 0; i < n; i++) yuv[i] = h->h8[i];
  return ICU_OK;
}

}  // extern "C"
