// Shared device helpers for the MI355X (gfx950) row-column beamforming kernels.
//
// Time of flight is computed on the fly.  Every per-(voxel, element) receive
// term and per-(voxel, angle) transmit term is formed in float64 and reduced to
// an integer sample, a float32 fraction and a float32 unit phasor; a tap then
// combines one receive and one transmit term with integer/float32 arithmetic
// only.  This keeps the sample index exact to ~1e-7 samples and the phase exact
// to ~1e-7 rad at any depth, where a float32 phase 2*pi*fd*tau would carry
// ~5e-4 rad of rounding at 15 MHz and 20 mm.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <cstddef>
#include <cstdint>

#include "rcabeam.h"

namespace rcb {

constexpr int kWave = 64;
// Receive terms of channels outside the dynamic aperture carry this index.
constexpr int kInactive = -(1 << 30);
// Sample indices are clamped to +/- kIndexLimit so that sums cannot overflow.
constexpr double kIndexLimit = 268435456.0;  // 2^28

struct Term {  // 16 bytes: one receive or transmit delay, split for exact taps
  int i;       // integer sample part (kInactive: channel outside the aperture)
  float f;     // fractional sample part in [0, 1)
  float c;     // cos(2 pi fd tau)
  float s;     // sin(2 pi fd tau)
};

struct Tap {  // 16 bytes: one combined (voxel, element, angle) tap
  int i0;     // first sample row, < 0 when the tap contributes nothing
  float w;    // interpolation weight toward row i0 + 1
  float c;    // phasor real part
  float s;    // phasor imaginary part
};

__device__ __forceinline__ Term split_delay(double samples, double cycles) {
  double fl = floor(samples);
  fl = fmin(fmax(fl, -kIndexLimit), kIndexLimit);
  float frac = static_cast<float>(samples - fl);
  if (frac >= 1.0f) {  // float rounding of a fraction just below one
    fl += 1.0;
    frac = 0.0f;
  }
  // Phase: the cycle count is reduced to [-1/2, 1/2] in float64, so the
  // float32 sincospi that follows sees a small argument (its error, ~1 ulp,
  // matches the float32 phasor it produces).
  const float cyc = static_cast<float>(cycles - rint(cycles));
  float sn, cs;
  sincospif(2.0f * cyc, &sn, &cs);
  return Term{static_cast<int>(fl), frac, cs, sn};
}

// Receive delay sqrt(z^2 + (v - e)^2) / c with the dynamic-aperture mask
// |v - e| <= z / (2 F); F <= 0 disables the mask.
__device__ __forceinline__ Term receive_term(double v, double z, double e, double c, double fs, double fd,
                                             double fn) {
  double dv = v - e;
  if (fn > 0.0 && !(fabs(dv) <= z / (2.0 * fn))) return Term{kInactive, 0.f, 0.f, 0.f};
  double tau = sqrt(z * z + dv * dv) / c;
  return split_delay(tau * fs, tau * fd);
}

// Plane-wave transmit delay (z cos a + u sin a) / c, sample origin t0.
__device__ __forceinline__ Term transmit_term(double u, double z, double cos_a, double sin_a, double t0, double c,
                                              double fs, double fd) {
  double tau = (z * cos_a + u * sin_a) / c;
  return split_delay((tau - t0) * fs, tau * fd);
}

// Tap validity follows the closed window 0 <= s <= ns - 1: at s == ns - 1 the
// value is row ns - 1 itself; outside, the tap is exactly zero.
__device__ __forceinline__ Tap combine(const Term& r, const Term& t, int ns) {
  Tap out{-1, 0.f, 0.f, 0.f};
  if (r.i == kInactive) return out;
  int i = r.i + t.i;
  float f = r.f + t.f;
  if (f >= 1.0f) {
    i += 1;
    f -= 1.0f;
  }
  bool valid = i >= 0 && (i < ns - 1 || (i == ns - 1 && f == 0.0f));
  if (!valid) return out;
  out.i0 = i;
  out.w = f;
  out.c = r.c * t.c - r.s * t.s;
  out.s = r.c * t.s + r.s * t.c;
  return out;
}

// ---- complex helpers -------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float cnorm(float2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ float cabs(float2 a) { return hypotf(a.x, a.y); }
// Signed square root x / sqrt(|x|), zero at zero.
__device__ __forceinline__ float2 sgnsqrt(float2 a) {
  float m = hypotf(a.x, a.y);
  if (!(m > 0.0f)) return make_float2(0.f, 0.f);
  float r = 1.0f / sqrtf(m);
  return make_float2(a.x * r, a.y * r);
}

// ---- sample storage --------------------------------------------------------
// Complex samples are stored as float2 (complex64), __half2 (FP16) or two
// packed bfloat16 values; arithmetic is always float32.
struct bf16x2 {
  uint32_t bits;  // low half: real, high half: imaginary
};

template <typename T>
__device__ __forceinline__ float2 load_sample(const T* p);

template <>
__device__ __forceinline__ float2 load_sample<float2>(const float2* p) {
  return *p;
}
template <>
__device__ __forceinline__ float2 load_sample<__half2>(const __half2* p) {
  return __half22float2(*p);
}
template <>
__device__ __forceinline__ float2 load_sample<bf16x2>(const bf16x2* p) {
  uint32_t b = p->bits;
  return make_float2(__uint_as_float(b << 16), __uint_as_float(b & 0xffff0000u));
}

// Two float32 values rounded to nearest even into one word of bfloat16 (lo
// in the low half): one gfx950 v_cvt_pk_bf16_f32.
typedef __bf16 bf16v2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t bf16_pack2_rne(float lo, float hi) {
  const bf16v2 v = {static_cast<__bf16>(lo), static_cast<__bf16>(hi)};
  return __builtin_bit_cast(uint32_t, v);
}

__device__ __forceinline__ void store_sample(float2* p, float2 v) { *p = v; }
__device__ __forceinline__ void store_sample(__half2* p, float2 v) { *p = __floats2half2_rn(v.x, v.y); }
__device__ __forceinline__ void store_sample(bf16x2* p, float2 v) { p->bits = bf16_pack2_rne(v.x, v.y); }

// Wave-wide sum (64 lanes) by butterfly shuffles; every lane gets the total.
__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) x += __shfl_xor(x, offset, kWave);
  return x;
}

// Split FP16 channel storage: each complex64 sample x is kept as two FP16
// pairs, hi = fp16(x 2^e) and lo = fp16(x 2^e - hi), with one power-of-two
// scale 2^e per ensemble chosen from the joint max |component| m so that
// m 2^e lies in [2^14, 2^15): hi + lo then carries 22 significant bits for
// every component within 2^17 of m and an absolute error below 2^-40 m for
// the rest.  absmax points at the bits of m (rcabeam_iq_absmax_device's
// out[0]); m = 0 or a non-finite m gives e = 0.
__device__ __forceinline__ int split_exponent(const unsigned* absmax) {
  const float m = __uint_as_float(absmax[0]);
  if (!(m > 0.f) || !(m < __builtin_inff())) return 0;
  int ex = 0;
  frexpf(m, &ex);  // m = f 2^ex, f in [0.5, 1)
  return 15 - ex;
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uniform(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

// Allow `kernel` up to `bytes` of dynamic LDS (beyond the 64 KiB default).
// The attribute belongs to the device current at the call, so it is set
// before every launch rather than cached (a host call of a few microseconds).
inline hipError_t allow_dynamic_lds(const void* kernel, int bytes) {
  return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

inline rcabeam_status launch_status() {
  return hipGetLastError() == hipSuccess ? RCABEAM_SUCCESS : RCABEAM_ERROR_CUDA;
}

}  // namespace rcb
