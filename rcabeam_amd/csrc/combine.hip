// Combiners on a grid from per-angle / per-channel grid launches for MI355X
// (gfx950): RC-FMAS, St-SW and DMAS-CCF-ACF.
//
// RC-FMAS and St-SW need, per voxel and frame, each angle's channel sum of each
// aperture (reference src/rcabeam/ensemble.cu:236-315 with METHOD 2 and
// 459-552).  On a grid, the LDS-tiled grid delay-and-sum produces exactly that
// when it runs once per angle in separate-aperture mode (method 5, one angle,
// channels packed angle-major).  These kernels fold each angle's (RC, CR) pair
// into running sums and reduce them over frames at the end:
//   * RC-FMAS: A = sum_a sgnsqrt(RC_a), B = sum_a sgnsqrt(CR_a); signal A B,
//     power mean |A B|^2 (reference fmas.py:31-51, ensemble.cu:261-275);
//   * St-SW (k interleaved subsets): R = sum_a RC_a, Q = sum_a CR_a and, per
//     subset s (angles a = s mod k), the sums of RC_a, CR_a, |RC_a|^2,
//     |CR_a|^2; the finish forms the subset coherences c[i, j] per frame, their
//     energies over frames, and |sum over pairs (i1 < i2, j1 != j2) of
//     <c[i1, j1], c[i2, j2]>| / norms as the weight, with |mean R conj(Q)| as
//     the power (reference stsw.py:10-82, ensemble.cu:459-590).
//
// Accumulator layout, float32 [voxel][field][frame]:
//   RC-FMAS fields 0-3: A.re, A.im, B.re, B.im
//   St-SW   fields 0-3: R.re, R.im, Q.re, Q.im; then per subset s, fields
//           4 + 6 s + (0..5): sum RC (re, im), sum CR (re, im), sum |RC|^2, sum |CR|^2
// The first angle that touches a field stores instead of adding, so the
// accumulators need no clearing.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "ensemble.hpp"
#include "rcabeam_mi355x.h"

namespace rcb {
namespace {

constexpr int kMaxSubsets = 4;

__host__ __device__ constexpr int accumulator_fields(int method, int k) { return method == kRcFmas ? 4 : 4 + 6 * k; }

constexpr int kAngleMaxBatch = 8;

// Fold `count` consecutive angles a0 .. a0 + count - 1 (their separate-aperture
// outputs `stride` complex64 apart) in one pass over the accumulator.
template <int METHOD, int K>
__global__ __launch_bounds__(256) void k_angle_accumulate(const float2* __restrict__ ap, long long stride,
                                                          int count, long long nv, int nt, int a0,
                                                          float* __restrict__ acc) {
  constexpr int F = accumulator_fields(METHOD, K);
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= nv * nt) return;
  const long long v = idx / nt;
  const int t = static_cast<int>(idx - v * nt);
  float* base = acc + v * F * nt + t;
  auto put = [&](int field, float value, bool first) {
    float* p = base + static_cast<long long>(field) * nt;
    *p = first ? value : *p + value;
  };
  float2 r_sum = make_float2(0.f, 0.f), q_sum = make_float2(0.f, 0.f);
  float2 mr[K], mq[K];
  float er[K], eq[K];
#pragma unroll
  for (int s = 0; s < K; ++s) {
    mr[s] = mq[s] = make_float2(0.f, 0.f);
    er[s] = eq[s] = 0.f;
  }
  for (int b = 0; b < count; ++b) {
    const float2* a = ap + b * stride;
    const float2 r = a[2 * v * nt + t];
    const float2 q = a[(2 * v + 1) * nt + t];
    if (METHOD == kRcFmas) {
      r_sum = cadd(r_sum, sgnsqrt(r));
      q_sum = cadd(q_sum, sgnsqrt(q));
    } else {
      r_sum = cadd(r_sum, r);
      q_sum = cadd(q_sum, q);
      const int sub = (a0 + b) % K;
#pragma unroll
      for (int s = 0; s < K; ++s)
        if (s == sub) {
          mr[s] = cadd(mr[s], r);
          mq[s] = cadd(mq[s], q);
          er[s] += cnorm(r);
          eq[s] += cnorm(q);
        }
    }
  }
  put(0, r_sum.x, a0 == 0);
  put(1, r_sum.y, a0 == 0);
  put(2, q_sum.x, a0 == 0);
  put(3, q_sum.y, a0 == 0);
  if (METHOD == kStsw) {
#pragma unroll
    for (int s = 0; s < K; ++s) {
      // Subset s holds the angles a = s mod K; its first is angle s.
      const bool touched = count >= K || ((s - a0 % K + K) % K) < count;
      if (!touched) continue;
      const bool first = s >= a0 && s < a0 + count;
      put(4 + 6 * s, mr[s].x, first);
      put(5 + 6 * s, mr[s].y, first);
      put(6 + 6 * s, mq[s].x, first);
      put(7 + 6 * s, mq[s].y, first);
      put(8 + 6 * s, er[s], first);
      put(9 + 6 * s, eq[s], first);
    }
  }
}

// One wave per voxel: signal = A B, power = mean |A B|^2.
__global__ __launch_bounds__(256) void k_fmas_finish(const float* __restrict__ acc, long long nv, int nt,
                                                     float* __restrict__ pd, float2* __restrict__ signal) {
  const long long v = static_cast<long long>(blockIdx.x) * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (v >= nv) return;
  const float* base = acc + v * 4 * nt;
  float pw = 0.f;
  for (int t = lane; t < nt; t += kWave) {
    const float2 a = make_float2(base[t], base[nt + t]);
    const float2 b = make_float2(base[2 * nt + t], base[3 * nt + t]);
    const float2 value = cmul(a, b);
    if (signal) signal[v * nt + t] = value;
    pw += cnorm(value);
  }
  pw = wave_sum(pw);
  if (lane == 0) {
    pd[2 * v] = pw / static_cast<float>(nt);
    pd[2 * v + 1] = 0.f;
  }
}

// Subset coherences of one frame (reference ensemble.cu:509-531).
template <int K>
__device__ __forceinline__ void coherences(const float* base, int nt, int t, int na, float2 (&c)[K * K]) {
  float2 mr[K], mq[K];
  float vr[K], vq[K];
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const float count = static_cast<float>((na - 1 - s) / K + 1);
    const float* f = base + static_cast<long long>(4 + 6 * s) * nt + t;
    mr[s] = make_float2(f[0] / count, f[nt] / count);
    mq[s] = make_float2(f[2 * nt] / count, f[3 * nt] / count);
    vr[s] = fmaxf(f[4 * nt] / count - cnorm(mr[s]), 0.f);
    vq[s] = fmaxf(f[5 * nt] / count - cnorm(mq[s]), 0.f);
  }
#pragma unroll
  for (int i = 0; i < K; ++i)
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const float2 num = cmulc(mr[i], mq[j]);
      const float den = cnorm(mr[i]) + cnorm(mq[j]) + 0.5f * (vr[i] + vq[j]) - cabs(num);
      float2 value = den > 0.f ? make_float2(num.x / den, num.y / den) : make_float2(0.f, 0.f);
      if (!isfinite(value.x) || !isfinite(value.y)) value = make_float2(0.f, 0.f);
      c[i * K + j] = value;
    }
}

// One wave per voxel: XDoppler power and the unnormalised St-SW weight.  Two
// passes over the frames (coherence energies, then normalised pair products);
// the pairs (i1 < i2, j1 != j2) are visited with running row / column sums.
template <int K>
__global__ __launch_bounds__(256) void k_stsw_finish(const float* __restrict__ acc, long long nv, int nt, int na,
                                                     float* __restrict__ pd, float* __restrict__ weights) {
  constexpr int F = accumulator_fields(kStsw, K);
  const long long v = static_cast<long long>(blockIdx.x) * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (v >= nv) return;
  const float* base = acc + v * F * nt;
  float energy[K * K];
#pragma unroll
  for (int s = 0; s < K * K; ++s) energy[s] = 0.f;
  float xr = 0.f, xi = 0.f;
  for (int t = lane; t < nt; t += kWave) {
    const float2 r = make_float2(base[t], base[nt + t]);
    const float2 q = make_float2(base[2 * nt + t], base[3 * nt + t]);
    const float2 cross = cmulc(r, q);
    xr += cross.x;
    xi += cross.y;
    float2 c[K * K];
    coherences<K>(base, nt, t, na, c);
#pragma unroll
    for (int s = 0; s < K * K; ++s) energy[s] += cnorm(c[s]);
  }
  float inv_norm[K * K];
#pragma unroll
  for (int s = 0; s < K * K; ++s) {
    const float e = wave_sum(energy[s]);
    inv_norm[s] = e > 0.f ? 1.0f / sqrtf(e) : 0.f;
  }
  float2 total_acc = make_float2(0.f, 0.f);
  for (int t = lane; t < nt; t += kWave) {
    float2 c[K * K];
    coherences<K>(base, nt, t, na, c);
    float2 above = make_float2(0.f, 0.f), column[K];
#pragma unroll
    for (int j = 0; j < K; ++j) column[j] = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < K; ++i) {
      float2 row = make_float2(0.f, 0.f);
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const float2 value = make_float2(c[i * K + j].x * inv_norm[i * K + j], c[i * K + j].y * inv_norm[i * K + j]);
        // Earlier rows, every column but j, times conj(value).
        const float2 earlier = make_float2(above.x - column[j].x, above.y - column[j].y);
        total_acc = cadd(total_acc, cmulc(earlier, value));
        column[j] = cadd(column[j], value);
        row = cadd(row, value);
      }
      above = cadd(above, row);
    }
  }
  xr = wave_sum(xr);
  xi = wave_sum(xi);
  total_acc.x = wave_sum(total_acc.x);
  total_acc.y = wave_sum(total_acc.y);
  if (lane == 0) {
    const float inv = 1.0f / static_cast<float>(nt);
    pd[2 * v] = hypotf(xr * inv, xi * inv);
    pd[2 * v + 1] = xi * inv;
    weights[v] = hypotf(total_acc.x, total_acc.y);
  }
}

// DMAS-CCF-ACF (reference src/rcabeam/dmas.py:32-85, ensemble.cu:143-234) from
// two kinds of grid launches: per channel i (both apertures' channel i, all
// angles: the angle-summed channel signals s) and per angle (each aperture's
// channel sum of one angle: the ACF's per-angle signals).  Fields, float32
// [voxel][field][frame]: 0-1 sum sgnsqrt(s), 2 sum |s|, 3-4 sum RC_a,
// 5 sum |RC_a|^2, 6-7 sum CR_a, 8 sum |CR_a|^2.
constexpr int kDmasFields = 9;

constexpr int kDmasMaxBatch = 8;

// PER_CHANNEL folds `count` separate-aperture outputs (one per channel, `stride`
// complex64 apart) in one pass over the accumulator.
template <bool PER_CHANNEL>
__global__ __launch_bounds__(256) void k_dmas_accumulate(const float2* __restrict__ ap, long long stride, int count,
                                                         long long nv, int nt, bool first, float* __restrict__ acc) {
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= nv * nt) return;
  const long long v = idx / nt;
  const int t = static_cast<int>(idx - v * nt);
  float* base = acc + v * kDmasFields * nt + t;
  auto put = [&](int field, float value) {
    float* p = base + static_cast<long long>(field) * nt;
    *p = first ? value : *p + value;
  };
  if (PER_CHANNEL) {
    float2 sum = make_float2(0.f, 0.f);
    float mag = 0.f;
    for (int b = 0; b < count; ++b) {
      const float2* a = ap + b * stride;
      const float2 r = a[2 * v * nt + t];
      const float2 q = a[(2 * v + 1) * nt + t];
      sum = cadd(sum, cadd(sgnsqrt(r), sgnsqrt(q)));
      mag += cabs(r) + cabs(q);
    }
    put(0, sum.x);
    put(1, sum.y);
    put(2, mag);
  } else {
    const float2 r = ap[2 * v * nt + t];
    const float2 q = ap[(2 * v + 1) * nt + t];
    put(3, r.x);
    put(4, r.y);
    put(5, cnorm(r));
    put(6, q.x);
    put(7, q.y);
    put(8, cnorm(q));
  }
}

// One wave per voxel of the grid slab (x, z, y): channels inside the dynamic
// aperture of both apertures (the CCF normaliser, float64 mask as the
// delay-and-sum kernels apply it), then per frame DMAS x CCF x ACF, averaged.
__global__ __launch_bounds__(256) void k_dmas_finish(const float* __restrict__ acc, long long nv, int nt, int na,
                                                     const float* __restrict__ x, const float* __restrict__ z,
                                                     const float* __restrict__ y, int nz, int ny,
                                                     const float* __restrict__ xe, const float* __restrict__ ye,
                                                     int nc, double fn, float* __restrict__ pd) {
  const long long v = static_cast<long long>(blockIdx.x) * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (v >= nv) return;
  const int iy = static_cast<int>(v % ny);
  const int iz = static_cast<int>((v / ny) % nz);
  const long long ix = v / ny / nz;
  const double vx = x[ix], vz = z[iz], vy = y[iy];
  float active = 0.f;
  for (int ch = lane; ch < nc; ch += kWave) {
    const double radius = vz / (2.0 * fn);
    active += (fn <= 0.0 || fabs(vx - static_cast<double>(xe[ch])) <= radius) ? 1.f : 0.f;
    active += (fn <= 0.0 || fabs(vy - static_cast<double>(ye[ch])) <= radius) ? 1.f : 0.f;
  }
  active = wave_sum(active);
  const float* base = acc + v * kDmasFields * nt;
  const float inv = 1.0f / static_cast<float>(na);
  float total = 0.f;
  for (int t = lane; t < nt; t += kWave) {
    const float2 sqrt_sum = make_float2(base[t], base[nt + t]);
    const float abs_sum = base[2 * nt + t];
    const float2 r = make_float2(base[3 * nt + t], base[4 * nt + t]);
    const float2 q = make_float2(base[6 * nt + t], base[7 * nt + t]);
    const float er = base[5 * nt + t], eq = base[8 * nt + t];
    const float dm = cnorm(sqrt_sum) - abs_sum;
    const float wc = (active > 0.f && abs_sum > 0.f) ? fmaxf(dm, 0.f) / active / abs_sum : 0.f;
    const float nr = cnorm(r) * inv * inv, nq = cnorm(q) * inv * inv;
    const float var = fmaxf(er * inv - nr, 0.f) + fmaxf(eq * inv - nq, 0.f);
    const float num = sqrtf(nr) * sqrtf(nq);
    const float den = nr + nq + 0.5f * var - num;
    float power = den > 0.f ? dm * wc * (num / den) : 0.f;
    if (!isfinite(power)) power = 0.f;
    total += power;
  }
  total = wave_sum(total);
  if (lane == 0) {
    pd[2 * v] = total / static_cast<float>(nt);
    pd[2 * v + 1] = 0.f;
  }
}

bool combine_args_ok(int method, int k) {
  if (method == kRcFmas) return true;
  return method == kStsw && k >= 2 && k <= kMaxSubsets;
}

}  // namespace
}  // namespace rcb

using namespace rcb;

extern "C" size_t rcabeam_angle_accumulator_fields(int method, int k) {
  return combine_args_ok(method, k) ? static_cast<size_t>(accumulator_fields(method, k)) : 0;
}

extern "C" rcabeam_status rcabeam_angle_accumulate_device(const void* apertures, size_t count, size_t stride,
                                                          size_t nv, size_t nt, int method, int k, int first_angle,
                                                          void* acc) {
  if (!apertures || !acc || nt < 1 || first_angle < 0 || count < 1 || count > kAngleMaxBatch ||
      (count > 1 && stride < 2 * nv * nt) || !combine_args_ok(method, k))
    return RCABEAM_ERROR_ARGUMENT;
  if (nv == 0) return RCABEAM_SUCCESS;
  const long long n = static_cast<long long>(nv) * static_cast<long long>(nt);
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  const float2* ap = static_cast<const float2*>(apertures);
  float* a = static_cast<float*>(acc);
  const long long lnv = static_cast<long long>(nv), ls = static_cast<long long>(stride);
  const int int_nt = static_cast<int>(nt), c = static_cast<int>(count);
  if (method == kRcFmas) {
    k_angle_accumulate<kRcFmas, 1><<<blocks, 256>>>(ap, ls, c, lnv, int_nt, first_angle, a);
  } else {
    switch (k) {
      case 2: k_angle_accumulate<kStsw, 2><<<blocks, 256>>>(ap, ls, c, lnv, int_nt, first_angle, a); break;
      case 3: k_angle_accumulate<kStsw, 3><<<blocks, 256>>>(ap, ls, c, lnv, int_nt, first_angle, a); break;
      default: k_angle_accumulate<kStsw, 4><<<blocks, 256>>>(ap, ls, c, lnv, int_nt, first_angle, a); break;
    }
  }
  return launch_status();
}

extern "C" rcabeam_status rcabeam_angle_combine_finish_device(const void* acc, size_t nv, size_t nt, size_t na,
                                                              int method, int k, float* pd, void* signal,
                                                              float* weights) {
  if (!acc || !pd || nt < 1 || na < 1 || !combine_args_ok(method, k)) return RCABEAM_ERROR_ARGUMENT;
  if (method == kStsw && (!weights || na < static_cast<size_t>(k))) return RCABEAM_ERROR_ARGUMENT;
  if (nv == 0) return RCABEAM_SUCCESS;
  const unsigned blocks = static_cast<unsigned>((nv + 3) / 4);
  const float* a = static_cast<const float*>(acc);
  const long long lnv = static_cast<long long>(nv);
  const int int_nt = static_cast<int>(nt), int_na = static_cast<int>(na);
  if (method == kRcFmas) {
    k_fmas_finish<<<blocks, 256>>>(a, lnv, int_nt, pd, static_cast<float2*>(signal));
  } else {
    switch (k) {
      case 2: k_stsw_finish<2><<<blocks, 256>>>(a, lnv, int_nt, int_na, pd, weights); break;
      case 3: k_stsw_finish<3><<<blocks, 256>>>(a, lnv, int_nt, int_na, pd, weights); break;
      default: k_stsw_finish<4><<<blocks, 256>>>(a, lnv, int_nt, int_na, pd, weights); break;
    }
  }
  return launch_status();
}

extern "C" size_t rcabeam_dmas_accumulator_fields(void) { return kDmasFields; }

extern "C" rcabeam_status rcabeam_dmas_accumulate_device(const void* apertures, size_t count, size_t stride,
                                                         size_t nv, size_t nt, int per_channel, int first,
                                                         void* acc) {
  if (!apertures || !acc || nt < 1 || count < 1 || count > kDmasMaxBatch || (!per_channel && count != 1) ||
      (count > 1 && stride < 2 * nv * nt))
    return RCABEAM_ERROR_ARGUMENT;
  if (nv == 0) return RCABEAM_SUCCESS;
  const long long n = static_cast<long long>(nv) * static_cast<long long>(nt);
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  const float2* ap = static_cast<const float2*>(apertures);
  float* a = static_cast<float*>(acc);
  const long long ls = static_cast<long long>(stride);
  const int c = static_cast<int>(count);
  if (per_channel)
    k_dmas_accumulate<true><<<blocks, 256>>>(ap, ls, c, static_cast<long long>(nv), static_cast<int>(nt), first != 0, a);
  else
    k_dmas_accumulate<false><<<blocks, 256>>>(ap, ls, c, static_cast<long long>(nv), static_cast<int>(nt), first != 0,
                                              a);
  return launch_status();
}

extern "C" rcabeam_status rcabeam_dmas_finish_device(const void* acc, const float* x, const float* z, const float* y,
                                                     size_t nx, size_t nz, size_t ny, const float* xe,
                                                     const float* ye, size_t nc, size_t na, size_t nt, float fn,
                                                     float* pd) {
  if (!acc || !x || !z || !y || !xe || !ye || !pd || nc < 1 || na < 1 || nt < 1 || std::isnan(fn))
    return RCABEAM_ERROR_ARGUMENT;
  const size_t nv = nx * nz * ny;
  if (nv == 0) return RCABEAM_SUCCESS;
  k_dmas_finish<<<static_cast<unsigned>((nv + 3) / 4), 256>>>(
      static_cast<const float*>(acc), static_cast<long long>(nv), static_cast<int>(nt), static_cast<int>(na), x, z, y,
      static_cast<int>(nz), static_cast<int>(ny), xe, ye, static_cast<int>(nc), static_cast<double>(fn), pd);
  return launch_status();
}
