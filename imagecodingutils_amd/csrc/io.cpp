// Raw YUV 4:2:0 file ingest / output and the motion-vector list dump
// (icu_yuv_file_*, icu_yuv_unpack_device, icu_mv_dump / icu_mv_load).
//
// Host side restates YuvRawFileHandler over RawFileHandlerBase
// (src/CodecUtils/YuvRawFileHandler.cpp:78-184,
// src/CodecUtils/RawFileHandlerBase.cpp:82-180) with the file mapped rather
// than copied; the device side unpacks a raw frame already in HBM into the
// 8-bit planes the estimator and compensator take.
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/icu_hip.h"
#include "icu_internal.h"

using namespace icu;

struct icu_yuv_file {
  std::string path;
  int rw = 0, W = 0, H = 0, in_type = 0, out_type = 0;
  // read: the mapped file
  const uint8_t* map = nullptr;
  long length = 0, pos = 0;
  // write: the stream buffer, written on close
  std::vector<uint8_t> wbuf;
  // the unit buffer; GetNextUnit overwrites only what it reads, so a short
  // last unit keeps the previous unit's tail
  std::vector<int16_t> b16;
  std::vector<uint8_t> b8;
  ~icu_yuv_file() {
    if (map) munmap((void*)map, (size_t)length);
  }
};

namespace {

int write_file(const std::string& path, const uint8_t* p, size_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return icu_fail(ICU_E_ARG, "cannot open " + path + " for writing");
  const size_t w = n ? std::fwrite(p, 1, n, f) : 0;
  std::fclose(f);
  return w == n ? ICU_OK : icu_fail(ICU_E_STATE, "short write to " + path);
}

}  // namespace

extern "C" {

int icu_yuv_file_open(const char* path, int rw, int width, int height, int in_type, int out_type,
                      icu_yuv_file** out) {
  if (!path || !out) return icu_fail(ICU_E_ARG, "null argument");
  if (rw != 0 && rw != 1) return icu_fail(ICU_E_ARG, "rw must be 0 (read) or 1 (write)");
  if (width < 2 || height < 2 || width % 2 || height % 2)
    return icu_fail(ICU_E_ARG, "width and height must be even and positive");
  if ((in_type != ICU_YUV420_8P && in_type != ICU_YUV420_NV12) ||
      (out_type != ICU_YUV420_8P && out_type != ICU_YUV420_16P))
    return icu_fail(ICU_E_ARG, "unknown YUV type");
  icu_yuv_file* f = new icu_yuv_file();
  f->path = path;
  f->rw = rw;
  f->W = width;
  f->H = height;
  f->in_type = in_type;
  f->out_type = out_type;
  const size_t n = (size_t)width * height * 3 / 2;
  if (out_type == ICU_YUV420_16P) f->b16.assign(n, 0);
  else f->b8.assign(n, 0);
  if (rw == 0) {
    const int fd = ::open(path, O_RDONLY);
    struct stat st;
    if (fd < 0 || fstat(fd, &st) != 0) {
      if (fd >= 0) ::close(fd);
      delete f;
      return icu_fail(ICU_E_ARG, std::string("cannot open ") + path);
    }
    f->length = (long)st.st_size;
    if (f->length > 0) {
      void* m = mmap(nullptr, (size_t)f->length, PROT_READ, MAP_PRIVATE, fd, 0);
      if (m == MAP_FAILED) {
        ::close(fd);
        delete f;
        return icu_fail(ICU_E_DEVICE, std::string("cannot map ") + path);
      }
      f->map = (const uint8_t*)m;
    }
    ::close(fd);
  }
  *out = f;
  return ICU_OK;
}

int icu_yuv_file_next_unit(icu_yuv_file* f, const void** unit, int* unit_len) {
  if (!f || !unit || !unit_len) return icu_fail(ICU_E_ARG, "null argument");
  if (f->rw != 0) return icu_fail(ICU_E_STATE, "file is open for writing");
  *unit = f->out_type == ICU_YUV420_16P ? (const void*)f->b16.data() : (const void*)f->b8.data();
  const long n = (long)f->W * f->H * 3 / 2, lum = (long)f->W * f->H, chr = lum / 4;
  long cnt = 0;
  if (f->pos >= f->length - 1) f->pos = 0;  // wrap around the end
  const uint8_t* src = f->map;
  if (f->in_type == ICU_YUV420_8P) {
    const long k = src ? std::min(n, f->length - f->pos) : 0;
    if (f->out_type == ICU_YUV420_16P)
      for (long i = 0; i < k; i++) f->b16[i] = src[f->pos + i];
    else
      std::memcpy(f->b8.data(), src + f->pos, (size_t)k);
    cnt = k;
    f->pos += k;
  } else if (f->out_type == ICU_YUV420_16P) {
    // NV12: planar luma, then interleaved U/V pairs split into the planes
    for (; cnt < lum && f->pos < f->length; cnt++) f->b16[cnt] = src[f->pos++];
    for (; cnt < lum + chr && f->pos < f->length; cnt++) {
      f->b16[cnt] = src[f->pos++];
      if (f->pos < f->length) f->b16[cnt + chr] = src[f->pos++];  // a file cut after U keeps the old V
    }
    // counts the V samples as written even when the file ended first
    cnt += chr;
  }
  // NV12 into 8-bit planes is not produced (unit length 0), as the handler
  *unit_len = (int)cnt;
  return ICU_OK;
}

int icu_yuv_file_put_unit(icu_yuv_file* f, const void* unit, int bytes) {
  if (!f || (!unit && bytes > 0) || bytes < 0) return icu_fail(ICU_E_ARG, "bad arguments");
  if (f->rw != 1) return icu_fail(ICU_E_STATE, "file is open for reading");
  const uint8_t* p = (const uint8_t*)unit;
  f->wbuf.insert(f->wbuf.end(), p, p + bytes);
  return ICU_OK;
}

int icu_yuv_file_close(icu_yuv_file* f) {
  if (!f) return icu_fail(ICU_E_ARG, "null argument");
  int rc = ICU_OK;
  // writing is delayed to close, and an empty stream writes no file
  if (f->rw == 1 && !f->wbuf.empty()) rc = write_file(f->path, f->wbuf.data(), f->wbuf.size());
  delete f;
  return rc;
}

int icu_mv_dump(const char* path, const char* title, const int16_t* mv, int n) {
  if (!path || (!mv && n > 0) || n < 0) return icu_fail(ICU_E_ARG, "bad arguments");
  (void)title;  // the reference sets its title before the table exists, so none is written
  std::string s = "mv x,mv y\n";
  char line[32];
  for (int i = 0; i < n; i++) {
    std::snprintf(line, sizeof(line), "%d,%d\n", (int)mv[2 * i], (int)mv[2 * i + 1]);
    s += line;
  }
  return write_file(path, (const uint8_t*)s.data(), s.size());
}

int icu_mv_load(const char* path, int16_t* mv, int max_n, int* n) {
  if (!path || !n || (!mv && max_n > 0)) return icu_fail(ICU_E_ARG, "null argument");
  FILE* f = std::fopen(path, "r");
  if (!f) return icu_fail(ICU_E_ARG, std::string("cannot open ") + path);
  char line[256];
  int k = 0, rc = ICU_OK;
  bool header = false;
  while (std::fgets(line, sizeof(line), f)) {
    int x, y;
    if (std::sscanf(line, "%d,%d", &x, &y) == 2) {
      if (k >= max_n) {
        rc = icu_fail(ICU_E_RANGE, "more vectors in the file than max_n");
        break;
      }
      if (x < -32768 || x > 32767 || y < -32768 || y > 32767) {
        rc = icu_fail(ICU_E_RANGE, "vector component outside int16");
        break;
      }
      mv[2 * k] = (int16_t)x;
      mv[2 * k + 1] = (int16_t)y;
      k++;
    } else if (!header && std::strncmp(line, "mv x", 4) == 0) {
      header = true;
    } else if (line[0] != '\n' && k > 0) {
      rc = icu_fail(ICU_E_ARG, "malformed vector line");
      break;
    }
  }
  std::fclose(f);
  *n = k;
  return rc;
}

}  // extern "C"

// ------------------------------------------------------------ device unpack
namespace {

// NV12 chroma: interleaved U/V pairs into the U and V planes, 4 pairs per lane.
__global__ __launch_bounds__(256) void nv12_split_kernel(const uint8_t* __restrict__ uv, long pairs,
                                                         uint8_t* __restrict__ u, uint8_t* __restrict__ v) {
  const long i = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  const bool aligned = ((((uintptr_t)uv) & 7) | (((uintptr_t)u | (uintptr_t)v) & 3)) == 0;
  if (i + 4 <= pairs && aligned) {
    const uint2 w = *(const uint2*)(uv + 2 * i);
    // bytes u0 v0 u1 v1 | u2 v2 u3 v3
    *(uint32_t*)(u + i) = __builtin_amdgcn_perm(w.y, w.x, 0x06040200u);
    *(uint32_t*)(v + i) = __builtin_amdgcn_perm(w.y, w.x, 0x07050301u);
  } else {
    for (long k = i; k < pairs && k < i + 4; k++) {
      u[k] = uv[2 * k];
      v[k] = uv[2 * k + 1];
    }
  }
}

}  // namespace

extern "C" int icu_yuv_unpack_device(const uint8_t* d_raw, int in_type, int width, int height, uint8_t* d_y,
                                     uint8_t* d_u, uint8_t* d_v, void* stream) {
  if (!d_raw || !d_y || !d_u || !d_v) return icu_fail(ICU_E_ARG, "null argument");
  if (width < 2 || height < 2 || width % 2 || height % 2)
    return icu_fail(ICU_E_ARG, "width and height must be even and positive");
  const hipStream_t st = (hipStream_t)stream;
  const size_t lum = (size_t)width * height, chr = lum / 4;
  if (hipMemcpyAsync(d_y, d_raw, lum, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return icu_fail(ICU_E_DEVICE, "luma copy failed");
  if (in_type == ICU_YUV420_8P) {
    if (hipMemcpyAsync(d_u, d_raw + lum, chr, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_v, d_raw + lum + chr, chr, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return icu_fail(ICU_E_DEVICE, "chroma copy failed");
  } else if (in_type == ICU_YUV420_NV12) {
    const long blocks = ((long)chr / 4 + 1 + 255) / 256;
    hipLaunchKernelGGL(nv12_split_kernel, dim3((unsigned)blocks), dim3(256), 0, st, d_raw + lum, (long)chr,
                       d_u, d_v);
    if (hipGetLastError() != hipSuccess) return icu_fail(ICU_E_DEVICE, "NV12 split launch failed");
  } else {
    return icu_fail(ICU_E_ARG, "unknown YUV type");
  }
  return ICU_OK;
}
