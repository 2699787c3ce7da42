// Channel packing: (samples, channels, angles, frames) -> (channels, angles,
// samples, frames), frames innermost, with complex64 -> {complex64, FP16, BF16}
// conversion (reference src/rcabeam/ensemble.cu:119-141, 630-655).
//
// The transpose moves whole frame rows: a (sample, channel, angle) row of nt
// complex values is contiguous in both layouts, so one wave copies one row with
// coalesced reads and writes (64 frames per instruction), no LDS tile needed.
// Range errors (FP16: non-finite or |component| > 65504; BF16: non-finite or
// beyond the largest finite bfloat16) are OR-reduced per wave and written once
// per wave with a plain store of 1 (idempotent), never per element.
#include "common.hpp"
#include "rcabeam_mi355x.h"

namespace rcb {

__device__ __forceinline__ void store_packed(float2* dst, float2 v) { *dst = v; }
__device__ __forceinline__ void store_packed(__half2* dst, float2 v) { *dst = __floats2half2_rn(v.x, v.y); }
__device__ __forceinline__ uint16_t bf16_bits_rn(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<uint16_t>((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ void store_packed(bf16x2* dst, float2 v) {
  dst->bits = static_cast<uint32_t>(bf16_bits_rn(v.x)) | (static_cast<uint32_t>(bf16_bits_rn(v.y)) << 16);
}

template <typename T>
__device__ __forceinline__ bool out_of_range(float2 v);
template <>
__device__ __forceinline__ bool out_of_range<float2>(float2) {
  return false;
}
template <>
__device__ __forceinline__ bool out_of_range<__half2>(float2 v) {
  return !(fabsf(v.x) <= 65504.f) || !(fabsf(v.y) <= 65504.f);  // also catches NaN
}
template <>
__device__ __forceinline__ bool out_of_range<bf16x2>(float2 v) {
  constexpr float kMaxBf16 = 3.3895313892515355e38f;  // largest finite bfloat16
  return !(fabsf(v.x) <= kMaxBf16) || !(fabsf(v.y) <= kMaxBf16);
}

// One wave per (channel, angle, sample) target row; rows are visited in
// target order so writes stream contiguously.
template <typename T>
__global__ __launch_bounds__(256) void pack_rows(const float2* __restrict__ src, T* __restrict__ dst,
                                                 int* range_error, int ns, int nc, int na, int nt,
                                                 long long rows) {
  const long long row = static_cast<long long>(blockIdx.x) * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= rows) return;
  // target row = (ch * na + a) * ns + s
  const long long s = row % ns;
  const long long ca = row / ns;
  const long long a = ca % na;
  const long long ch = ca / na;
  const float2* in = src + ((s * nc + ch) * na + a) * nt;
  T* out = dst + row * nt;
  bool bad = false;
  for (int t = lane; t < nt; t += kWave) {
    const float2 v = in[t];
    bad |= out_of_range<T>(v);
    store_packed(out + t, v);
  }
  if (range_error && __any(bad) && lane == 0) *range_error = 1;
}

// Split FP16 packing: one wave per target row; each sample becomes the word
// pair (hi, lo) of FP16 complex values (see split_exponent in common.hpp).
__global__ __launch_bounds__(256) void pack_rows_split(const float2* __restrict__ src, uint2* __restrict__ dst,
                                                       const unsigned* __restrict__ absmax, int ns, int nc, int na,
                                                       int nt, long long rows) {
  const long long row = static_cast<long long>(blockIdx.x) * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= rows) return;
  const int e = split_exponent(absmax);
  const long long s = row % ns;
  const long long ca = row / ns;
  const long long a = ca % na;
  const long long ch = ca / na;
  const float2* in = src + ((s * nc + ch) * na + a) * nt;
  uint2* out = dst + row * nt;
  for (int t = lane; t < nt; t += kWave) {
    const float2 v = in[t];
    const float xr = ldexpf(v.x, e), xi = ldexpf(v.y, e);
    const __half2 hi = __floats2half2_rn(xr, xi);
    const float2 hf = __half22float2(hi);
    const __half2 lo = __floats2half2_rn(xr - hf.x, xi - hf.y);
    out[t] = make_uint2(__builtin_bit_cast(uint32_t, hi), __builtin_bit_cast(uint32_t, lo));
  }
}

// Inverse of pack_rows for complex64: one wave per packed (channel, angle,
// sample) row, written back to (samples, channels, angles, frames).
__global__ __launch_bounds__(256) void unpack_rows(const float2* __restrict__ src, float2* __restrict__ dst, int ns,
                                                   int nc, int na, int nt, long long rows) {
  const long long row = static_cast<long long>(blockIdx.x) * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= rows) return;
  const long long s = row % ns;
  const long long ca = row / ns;
  const long long a = ca % na;
  const long long ch = ca / na;
  const float2* in = src + row * nt;
  float2* out = dst + ((s * nc + ch) * na + a) * nt;
  for (int t = lane; t < nt; t += kWave) out[t] = in[t];
}

template <typename T>
static rcabeam_status launch_pack(const void* source, void* target, int* range_error, size_t ns, size_t nc,
                                  size_t na, size_t nt) {
  const long long rows = static_cast<long long>(ns) * nc * na;
  const long long blocks = (rows + 3) / 4;
  if (blocks > 0x7fffffffLL) return RCABEAM_ERROR_ARGUMENT;
  pack_rows<T><<<dim3(static_cast<unsigned>(blocks)), 256>>>(static_cast<const float2*>(source),
                                                              static_cast<T*>(target), range_error,
                                                              static_cast<int>(ns), static_cast<int>(nc),
                                                              static_cast<int>(na), static_cast<int>(nt), rows);
  return launch_status();
}

static bool valid_dims(size_t ns, size_t nc, size_t na, size_t nt) {
  if (!ns || !nc || !na || !nt) return false;
  if (ns > 0x7fffffff || nc > 0x7fffffff || na > 0x7fffffff || nt > 0x7fffffff) return false;
  return ns <= SIZE_MAX / nc / na / nt / 8;
}

}  // namespace rcb

using namespace rcb;

extern "C" rcabeam_status rcabeam_pack_ensemble_device(const void* source, void* target, size_t ns, size_t nc,
                                                       size_t na, size_t nt) {
  if (!source || !target || !valid_dims(ns, nc, na, nt)) return RCABEAM_ERROR_ARGUMENT;
  return launch_pack<float2>(source, target, nullptr, ns, nc, na, nt);
}

extern "C" rcabeam_status rcabeam_pack_ensemble_fp16_device(const void* source, void* target, int* range_error,
                                                            size_t ns, size_t nc, size_t na, size_t nt) {
  if (!source || !target || !range_error || !valid_dims(ns, nc, na, nt) ||
      reinterpret_cast<uintptr_t>(source) % 8 || reinterpret_cast<uintptr_t>(target) % 4 ||
      reinterpret_cast<uintptr_t>(range_error) % 4)
    return RCABEAM_ERROR_ARGUMENT;
  return launch_pack<__half2>(source, target, range_error, ns, nc, na, nt);
}

extern "C" rcabeam_status rcabeam_pack_ensemble_bf16_device(const void* source, void* target, int* range_error,
                                                            size_t ns, size_t nc, size_t na, size_t nt) {
  if (!source || !target || !range_error || !valid_dims(ns, nc, na, nt) ||
      reinterpret_cast<uintptr_t>(source) % 8 || reinterpret_cast<uintptr_t>(target) % 4 ||
      reinterpret_cast<uintptr_t>(range_error) % 4)
    return RCABEAM_ERROR_ARGUMENT;
  return launch_pack<bf16x2>(source, target, range_error, ns, nc, na, nt);
}

extern "C" rcabeam_status rcabeam_unpack_ensemble_device(const void* source, void* target, size_t ns, size_t nc,
                                                         size_t na, size_t nt) {
  if (!source || !target || !valid_dims(ns, nc, na, nt)) return RCABEAM_ERROR_ARGUMENT;
  const long long rows = static_cast<long long>(ns) * nc * na;
  const long long blocks = (rows + 3) / 4;
  if (blocks > 0x7fffffffLL) return RCABEAM_ERROR_ARGUMENT;
  unpack_rows<<<dim3(static_cast<unsigned>(blocks)), 256>>>(static_cast<const float2*>(source),
                                                             static_cast<float2*>(target), static_cast<int>(ns),
                                                             static_cast<int>(nc), static_cast<int>(na),
                                                             static_cast<int>(nt), rows);
  return launch_status();
}

extern "C" rcabeam_status rcabeam_pack_ensemble_split_device(const void* source, void* target, const unsigned* absmax,
                                                             size_t ns, size_t nc, size_t na, size_t nt) {
  if (!source || !target || !absmax || !valid_dims(ns, nc, na, nt) || reinterpret_cast<uintptr_t>(source) % 8 ||
      reinterpret_cast<uintptr_t>(target) % 8 || reinterpret_cast<uintptr_t>(absmax) % 4)
    return RCABEAM_ERROR_ARGUMENT;
  const long long rows = static_cast<long long>(ns) * nc * na;
  const long long blocks = (rows + 3) / 4;
  if (blocks > 0x7fffffffLL) return RCABEAM_ERROR_ARGUMENT;
  pack_rows_split<<<dim3(static_cast<unsigned>(blocks)), 256>>>(static_cast<const float2*>(source),
                                                                 static_cast<uint2*>(target), absmax,
                                                                 static_cast<int>(ns), static_cast<int>(nc),
                                                                 static_cast<int>(na), static_cast<int>(nt), rows);
  return launch_status();
}
