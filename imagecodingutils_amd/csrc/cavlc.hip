// CAVLC entropy coding of quantised 4x4 / 2x2 / 8x8 blocks (icu_cavlc_*):
// single blocks on the host, as CAVLCH264Impl::Encode / Decode
// (src/CodecUtils/CAVLCH264Impl.cpp:134-577) over an MSB-first bit stream
// (BitStreamWriterMSB / BitStreamReaderMSB), and whole batches on the device.
//
// Device encode: a counting pass (one lane per block, the encoder with a
// sink that only counts), an exclusive scan of the counts into bit offsets,
// then a writing pass in which each lane encodes its block again straight to
// its offset in one shared stream.  Blocks meet at arbitrary bit positions,
// so a lane ORs its bits into the (zeroed) stream 32 at a time with atomics;
// no lane ever writes a bit another lane owns.  Device decode takes the
// offsets the encoder produced and decodes every block in parallel.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>

#include "../../include/icu_hip.h"
#include "icu_cavlc.h"
#include "icu_internal.h"

using namespace icu;

namespace {

// ------------------------------------------------------------- bit access
// Host writer over a byte stream: sets or clears each bit (as the reference
// writer does), MSB of each byte first.
struct HostSink {
  uint8_t* s;
  long cap, pos;
  bool fits(int n) const { return pos + n <= cap; }
  bool put(uint32_t v, int n) {
    if (pos + n > cap) return false;
    if (s)
      for (int i = n - 1; i >= 0; i--, pos++) {
        const uint8_t m = (uint8_t)(0x80u >> (pos & 7));
        if ((v >> i) & 1u) s[pos >> 3] |= m;
        else s[pos >> 3] &= (uint8_t)~m;
      }
    else
      pos += n;
    return true;
  }
};

struct HostSource {
  const uint8_t* s;
  long cap, pos;
  __host__ uint32_t peek(int n) const {
    uint32_t v = 0;
    for (int i = 0; i < n; i++) {
      const long b = pos + i;
      v = (v << 1) | (b < cap ? (uint32_t)((s[b >> 3] >> (7 - (b & 7))) & 1u) : 0u);
    }
    return v;
  }
  __host__ void skip(int n) { pos += n; }
  __host__ long left() const { return cap - pos; }
};

struct CountSink {
  __device__ bool fits(int) const { return true; }
  __device__ bool put(uint32_t, int) { return true; }
};

// ORs bits into a zeroed stream of little-endian dwords read as an
// MSB-first byte stream: bit k is bit 7 - k % 8 of byte k / 8.
struct AtomicSink {
  uint32_t* s;
  long long pos;
  __device__ static uint32_t bswap(uint32_t x) { return __builtin_bswap32(x); }
  __device__ bool fits(int) const { return true; }  // the stream was sized by the counting pass
  __device__ bool put(uint32_t v, int n) {
    if (n == 0) return true;
    const uint32_t x = n == 32 ? v : (v & ((1u << n) - 1u));
    const long long q = pos >> 5;
    const int r = (int)(pos & 31);
    // the n bits at positions r .. r + n - 1 of a 64-bit big-endian window
    const uint64_t w = (uint64_t)x << (64 - r - n);
    const uint32_t hi = (uint32_t)(w >> 32), lo = (uint32_t)w;
    if (hi) atomicOr(s + q, bswap(hi));
    if (lo) atomicOr(s + q + 1, bswap(lo));
    pos += n;
    return true;
  }
};

struct DeviceSource {
  const uint32_t* s;
  long long ndw, cap, pos;
  __device__ uint32_t word(long long q) const {
    return (unsigned long long)q < (unsigned long long)ndw ? __builtin_bswap32(s[q]) : 0u;
  }
  __device__ uint32_t peek(int n) const {
    if (n == 0) return 0;
    const long long q = pos >> 5;
    const int r = (int)(pos & 31);
    const uint64_t w = ((uint64_t)word(q) << 32) | word(q + 1);
    uint32_t v = (uint32_t)((w << r) >> (64 - n));
    // bits past the end of the stream read as zeros
    const long long rem = cap - pos;
    if (rem < n) v = rem <= 0 ? 0u : (v & ~((1u << (n - (int)rem)) - 1u));
    return v;
  }
  __device__ void skip(int n) { pos += n; }
  __device__ long long left() const { return cap - pos; }
};

__device__ __forceinline__ int block_nc(const int8_t* d_nc, int nc, long b) { return d_nc ? d_nc[b] : nc; }

__global__ __launch_bounds__(256) void cavlc_count_kernel(const int16_t* __restrict__ coef, long n, int mode,
                                                          const int8_t* __restrict__ d_nc, int nc, int dc_skip,
                                                          long long* __restrict__ bits, int* __restrict__ err) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > n) return;
  if (b == n) {
    bits[n] = 0;  // so that the exclusive scan's last entry is the total
    return;
  }
  const int size = cavlc::block_size(mode);
  CountSink sink;
  const int r = cavlc::encode_block(coef + b * size, mode, block_nc(d_nc, nc, b), dc_skip, sink);
  if (r < 0) {
    atomicOr(err, 1);
    bits[b] = 0;
  } else {
    bits[b] = r;
  }
}

__global__ __launch_bounds__(256) void cavlc_write_kernel(const int16_t* __restrict__ coef, long n, int mode,
                                                          const int8_t* __restrict__ d_nc, int nc, int dc_skip,
                                                          const long long* __restrict__ offsets,
                                                          uint32_t* __restrict__ stream) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const int size = cavlc::block_size(mode);
  AtomicSink sink{stream, offsets[b]};
  cavlc::encode_block(coef + b * size, mode, block_nc(d_nc, nc, b), dc_skip, sink);
}

// MAXC: the largest block the launch decodes (16 for 2x2 / 4x4, 64 for 8x8)
template <int MAXC>
__global__ __launch_bounds__(256) void cavlc_decode_kernel(const uint32_t* __restrict__ stream, long long ndw,
                                                           long long cap_bits, const long long* __restrict__ offsets,
                                                           long n, int mode, const int8_t* __restrict__ d_nc, int nc,
                                                           int dc_skip, int16_t* __restrict__ coef,
                                                           int32_t* __restrict__ status) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const int size = cavlc::block_size(mode);
  DeviceSource src{stream, ndw, cap_bits, offsets[b]};
  int16_t out[MAXC];
  const int r = cavlc::decode_block(src, mode, block_nc(d_nc, nc, b), dc_skip, out);
  for (int i = 0; i < size; i++) coef[b * size + i] = out[i];
  if (status) status[b] = r;
}

__global__ __launch_bounds__(256) void zero_u32_kernel(uint32_t* p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0u;
}

int check_mode(int mode, int nc, int dc_skip) {
  if (mode != cavlc::kMode2x2 && mode != cavlc::kMode4x4 && mode != cavlc::kMode8x8)
    return icu_fail(ICU_E_ARG, "mode must be 0 (2x2), 1 (4x4) or 2 (8x8)");
  if (dc_skip != 0 && dc_skip != 1) return icu_fail(ICU_E_ARG, "dc_skip must be 0 or 1");
  if (nc < -2 || (mode != cavlc::kMode2x2 && nc < 0))
    return icu_fail(ICU_E_ARG, "nC must be >= 0 (4x4 and 8x8 blocks) or -1 / -2 / >= 0 (2x2 blocks)");
  return ICU_OK;
}

int err_code(int r) {
  return r == cavlc::kErrStream ? icu_fail(ICU_E_RANGE, "bit stream full or exhausted")
                                : icu_fail(ICU_E_STATE, "no such CAVLC symbol");
}

}  // namespace

struct icu_cavlc {
  void* temp = nullptr;
  size_t temp_bytes = 0;
  int* d_err = nullptr;
  ~icu_cavlc() {
    hipFree(temp);
    hipFree(d_err);
  }
};

extern "C" {

int icu_cavlc_encode(const int16_t* coef, int mode, int nc, int dc_skip, uint8_t* stream, long stream_bits,
                     long* bit_pos, int* nbits) {
  if (!coef || !bit_pos || !nbits) return icu_fail(ICU_E_ARG, "null argument");
  int rc = check_mode(mode, nc, dc_skip);
  if (rc) return rc;
  HostSink sink{stream, stream ? stream_bits : (1L << 62), *bit_pos};
  const int r = cavlc::encode_block(coef, mode, nc, dc_skip, sink);
  // on an error the symbols written before it stay, and so does the position
  *nbits = (int)(sink.pos - *bit_pos);
  *bit_pos = sink.pos;
  if (r < 0) return err_code(r);
  *nbits = r;
  return ICU_OK;
}

int icu_cavlc_decode(const uint8_t* stream, long stream_bits, long* bit_pos, int mode, int nc, int dc_skip,
                     int16_t* coef, int* nbits) {
  if (!stream || !bit_pos || !coef || !nbits) return icu_fail(ICU_E_ARG, "null argument");
  int rc = check_mode(mode, nc, dc_skip);
  if (rc) return rc;
  HostSource src{stream, stream_bits, *bit_pos};
  const int r = cavlc::decode_block(src, mode, nc, dc_skip, coef);
  if (r < 0) return err_code(r);
  *bit_pos = src.pos;
  *nbits = r;
  return ICU_OK;
}

int icu_cavlc_create(icu_cavlc** out) {
  if (!out) return icu_fail(ICU_E_ARG, "null argument");
  icu_cavlc* h = new icu_cavlc();
  if (hipMalloc(&h->d_err, sizeof(int)) != hipSuccess) {
    delete h;
    return icu_fail(ICU_E_DEVICE, "device allocation failed");
  }
  *out = h;
  return ICU_OK;
}

int icu_cavlc_destroy(icu_cavlc* h) {
  delete h;
  return ICU_OK;
}

int icu_cavlc_encode_device(icu_cavlc* h, const int16_t* d_coef, long nblocks, int mode, const int8_t* d_nc,
                            int nc, int dc_skip, uint8_t* d_stream, long stream_bytes, int64_t* d_offsets,
                            long long* total_bits, void* stream) {
  if (!h || !d_coef || !d_stream || !d_offsets || !total_bits || nblocks < 0)
    return icu_fail(ICU_E_ARG, "bad arguments");
  int rc = check_mode(mode, d_nc ? 0 : nc, dc_skip);
  if (rc) return rc;
  if (((uintptr_t)d_stream & 3) || (stream_bytes & 3))
    return icu_fail(ICU_E_ARG, "d_stream and stream_bytes must be multiples of 4");
  const hipStream_t st = (hipStream_t)stream;
  long long* offs = (long long*)d_offsets;
  const unsigned g1 = (unsigned)((nblocks + 1 + 255) / 256);
  if (hipMemsetAsync(h->d_err, 0, sizeof(int), st) != hipSuccess) return icu_fail(ICU_E_DEVICE, "memset failed");
  hipLaunchKernelGGL(cavlc_count_kernel, dim3(g1), dim3(256), 0, st, d_coef, nblocks, mode, d_nc, nc, dc_skip,
                     offs, h->d_err);
  size_t need = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, need, offs, offs, (int)(nblocks + 1), st) != hipSuccess)
    return icu_fail(ICU_E_DEVICE, "scan sizing failed");
  if (need > h->temp_bytes) {
    hipFree(h->temp);
    h->temp = nullptr;
    h->temp_bytes = 0;
    if (hipMalloc(&h->temp, need) != hipSuccess) return icu_fail(ICU_E_DEVICE, "scan storage allocation failed");
    h->temp_bytes = need;
  }
  size_t have = h->temp_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(h->temp, have, offs, offs, (int)(nblocks + 1), st) != hipSuccess)
    return icu_fail(ICU_E_DEVICE, "scan failed");
  long long tot = 0;
  int err = 0;
  if (hipMemcpyAsync(&tot, offs + nblocks, sizeof(tot), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(&err, h->d_err, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return icu_fail(ICU_E_DEVICE, "reading the stream length failed");
  if (err) return icu_fail(ICU_E_STATE, "a block has no CAVLC symbol for its nC / mode");
  *total_bits = tot;
  if (tot > (long long)stream_bytes * 8) return icu_fail(ICU_E_RANGE, "the stream does not fit d_stream");
  const long long ndw = (tot + 31) / 32;
  if (ndw > 0)
    hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)std::min<long long>(2048, (ndw + 255) / 256)), dim3(256), 0,
                       st, (uint32_t*)d_stream, ndw);
  if (nblocks > 0)
    hipLaunchKernelGGL(cavlc_write_kernel, dim3((unsigned)((nblocks + 255) / 256)), dim3(256), 0, st, d_coef,
                       nblocks, mode, d_nc, nc, dc_skip, (const long long*)offs, (uint32_t*)d_stream);
  if (hipGetLastError() != hipSuccess) return icu_fail(ICU_E_DEVICE, "CAVLC launch failed");
  return ICU_OK;
}

int icu_cavlc_decode_device(icu_cavlc* h, const uint8_t* d_stream, long stream_bytes, const int64_t* d_offsets,
                            long nblocks, int mode, const int8_t* d_nc, int nc, int dc_skip, int16_t* d_coef,
                            int32_t* d_status, void* stream) {
  if (!h || !d_stream || !d_offsets || !d_coef || nblocks < 0) return icu_fail(ICU_E_ARG, "bad arguments");
  int rc = check_mode(mode, d_nc ? 0 : nc, dc_skip);
  if (rc) return rc;
  if (((uintptr_t)d_stream & 3) || (stream_bytes & 3))
    return icu_fail(ICU_E_ARG, "d_stream and stream_bytes must be multiples of 4");
  if (nblocks > 0)
    hipLaunchKernelGGL(mode == cavlc::kMode8x8 ? cavlc_decode_kernel<64> : cavlc_decode_kernel<16>,
                       dim3((unsigned)((nblocks + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const uint32_t*)d_stream, (long long)stream_bytes / 4, (long long)stream_bytes * 8,
                       (const long long*)d_offsets, nblocks, mode, d_nc, nc, dc_skip, d_coef, d_status);
  if (hipGetLastError() != hipSuccess) return icu_fail(ICU_E_DEVICE, "CAVLC launch failed");
  return ICU_OK;
}

}  // extern "C"
