// Low-rank + residual coding of 16-bit channel storage for the mixed-precision
// OPW (no reference counterpart: the reference's FP16 path rounds the channels
// directly, src/rcabeam/ensemble.cu:129-141).
//
// Quantizing channel data to 16 bits rounds the strong tissue echoes with a
// relative error of 2^-12 (FP16) or 2^-9 (BF16).  That error changes from
// frame to frame, so the temporal clutter filter cannot remove it, and it
// lands on the weak (-40 dB) blood signal the filter keeps.  Coding removes
// the tissue before rounding.  With V (nt x k, orthonormal columns) spanning
// the strongest temporal modes of the channel Casorati matrix C (rows =
// (sample, channel, angle), columns = frames, both apertures together), every
// row splits exactly as
//     c = w V^H + e,   w = c V (k coefficients),   e = c - w V^H,
// and e (blood, noise, the clutter V misses) is what is rounded to 16 bits.
// The k coefficients are appended to the row as 2k extra "frames" (16-bit hi
// and lo parts of each w_j), so the packed row holds nt + 2k frames and the
// delay-and-sum treats them like frames.  Delay-and-sum D is linear and acts
// on every frame column alike, so its output on the coded rows gives
//     D(c)[t] = D(e)[t] + sum_j (D(w_hi)_j + D(w_lo)_j) conj(V[t, j]),
// which rcabeam_lowrank_expand_device forms.  Any V gives the exact result up
// to the rounding of e and of w; a V that captures the clutter makes e small.
// (Tap-coefficient rounding is the same for every frame column, so it maps the
// clutter's temporal subspace onto itself and the filter still removes it.)
#include "common.hpp"
#include "rcabeam_mi355x.h"
#include "tile_staging.hpp"

namespace rcb {
namespace {

constexpr int kMaxBasis = 12;        // k <= 12 keeps nt + 2k <= 224 for nt = 200
constexpr int kMaxFramesLR = 256;    // 4 frames per lane

template <typename T>
__device__ __forceinline__ bool in_range(float v);
template <>
__device__ __forceinline__ bool in_range<__half2>(float v) {
  return fabsf(v) <= 65504.f;  // false for NaN
}
template <>
__device__ __forceinline__ bool in_range<bf16x2>(float v) {
  return fabsf(v) <= 3.3895313892515355e38f;
}

// One wave per target (channel, angle, sample) row: frames t = lane + 64 j.
// Coefficients and residual in float64 (the row is read once; HBM-bound).
template <typename T>
__global__ __launch_bounds__(256) void k_pack_lowrank(const float2* __restrict__ src, uint32_t* __restrict__ dst,
                                                      const double2* __restrict__ basis, int k, int* range_error,
                                                      int ns, int nc, int na, int nt, long long rows) {
  const long long row = static_cast<long long>(blockIdx.x) * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= rows) return;
  const long long s = row % ns;
  const long long ca = row / ns;
  const long long a = ca % na;
  const long long ch = ca / na;
  const float2* in = src + ((s * nc + ch) * na + a) * nt;
  const int ntc = nt + 2 * k;
  uint32_t* out = dst + row * ntc;
  double2 x[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int t = lane + kWave * j;
    const float2 v = t < nt ? in[t] : make_float2(0.f, 0.f);
    x[j] = make_double2(v.x, v.y);
  }
  double2 e[4] = {x[0], x[1], x[2], x[3]};
  double2 w_mine = make_double2(0.0, 0.0);  // lane j < k keeps w_j
  for (int q = 0; q < k; ++q) {
    const double2* vq = basis + static_cast<long long>(q) * nt;
    double cr = 0.0, ci = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int t = lane + kWave * j;
      if (t < nt) {
        const double2 b = vq[t];  // w_q = sum_t c_t V[t, q]
        cr += x[j].x * b.x - x[j].y * b.y;
        ci += x[j].x * b.y + x[j].y * b.x;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      cr += __shfl_xor(cr, off, kWave);
      ci += __shfl_xor(ci, off, kWave);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int t = lane + kWave * j;
      if (t < nt) {
        const double2 b = vq[t];  // e_t -= w_q conj(V[t, q])
        e[j].x -= cr * b.x + ci * b.y;
        e[j].y -= ci * b.x - cr * b.y;
      }
    }
    if (lane == q) w_mine = make_double2(cr, ci);
  }
  bool bad = false;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int t = lane + kWave * j;
    if (t < nt) {
      const float re = static_cast<float>(e[j].x), im = static_cast<float>(e[j].y);
      bad |= !in_range<T>(re) || !in_range<T>(im);
      out[t] = pack2<T>(re, im);
    }
  }
  if (lane < k) {  // frames nt + 2 lane (hi) and nt + 2 lane + 1 (lo) hold w_lane
    const float re = static_cast<float>(w_mine.x), im = static_cast<float>(w_mine.y);
    bad |= !in_range<T>(re) || !in_range<T>(im);
    const uint32_t hi = pack2<T>(re, im);
    const float2 h = widen<T>(hi);
    const uint32_t lo = pack2<T>(static_cast<float>(w_mine.x - h.x), static_cast<float>(w_mine.y - h.y));
    out[nt + 2 * lane] = hi;
    out[nt + 2 * lane + 1] = lo;
  }
  if (range_error && __any(bad) && lane == 0) *range_error = 1;
}

// One wave per voxel row of the coded delay-and-sum output (nt + 2k complex64
// columns): x_t = d_t + sum_j (d[nt + 2j] + d[nt + 2j + 1]) conj(V[t, j]),
// stored in `O`; power (mean over the nt frames of |x_t|^2) into pd[2v].
template <typename O>
__global__ __launch_bounds__(256) void k_expand(const float2* __restrict__ coded, const double2* __restrict__ basis,
                                                int k, O* __restrict__ out, float* __restrict__ pd, int nt,
                                                long long rows) {
  const long long row = static_cast<long long>(blockIdx.x) * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= rows) return;
  const int ntc = nt + 2 * k;
  const float2* in = coded + row * ntc;
  // Lane j < k: the delay-and-sum of w_j (hi + lo parts).
  float2 dw = make_float2(0.f, 0.f);
  if (lane < k) {
    const float2 h = in[nt + 2 * lane], l = in[nt + 2 * lane + 1];
    dw = make_float2(h.x + l.x, h.y + l.y);
  }
  float2 x[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int t = lane + kWave * j;
    x[j] = t < nt ? in[t] : make_float2(0.f, 0.f);
  }
  for (int q = 0; q < k; ++q) {
    const float wr = __shfl(dw.x, q, kWave), wi = __shfl(dw.y, q, kWave);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int t = lane + kWave * j;
      if (t < nt) {
        const double2 b = basis[static_cast<long long>(q) * nt + t];
        const float br = static_cast<float>(b.x), bi = static_cast<float>(b.y);
        x[j].x = fmaf(wr, br, fmaf(wi, bi, x[j].x));  // + w conj(b)
        x[j].y = fmaf(wi, br, fmaf(-wr, bi, x[j].y));
      }
    }
  }
  float power = 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int t = lane + kWave * j;
    if (t < nt) {
      power = fmaf(x[j].x, x[j].x, fmaf(x[j].y, x[j].y, power));
      if (out) store_sample(out + row * nt + t, x[j]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) power += __shfl_xor(power, off, kWave);
  if (pd && lane == 0) {
    pd[2 * row] = power / static_cast<float>(nt);
    pd[2 * row + 1] = 0.f;
  }
}

}  // namespace
}  // namespace rcb

using namespace rcb;

extern "C" rcabeam_status rcabeam_pack_ensemble_lowrank_device(const void* source, void* target, int storage,
                                                               const void* basis, size_t k, int* range_error,
                                                               size_t ns, size_t nc, size_t na, size_t nt) {
  if (!source || !target || !range_error || (k && !basis) || !ns || !nc || !na || !nt) return RCABEAM_ERROR_ARGUMENT;
  if (k > kMaxBasis || nt > kMaxFramesLR || ns * nc * na > (1ull << 40)) return RCABEAM_ERROR_ARGUMENT;
  if (storage != RCABEAM_STORAGE_FP16 && storage != RCABEAM_STORAGE_BF16) return RCABEAM_ERROR_ARGUMENT;
  const long long rows = static_cast<long long>(ns * nc * na);
  const unsigned grid = static_cast<unsigned>((rows + 3) / 4);
  const double2* v = static_cast<const double2*>(basis);
  if (storage == RCABEAM_STORAGE_FP16)
    k_pack_lowrank<__half2><<<grid, 256>>>(static_cast<const float2*>(source), static_cast<uint32_t*>(target), v,
                                           static_cast<int>(k), range_error, static_cast<int>(ns),
                                           static_cast<int>(nc), static_cast<int>(na), static_cast<int>(nt), rows);
  else
    k_pack_lowrank<bf16x2><<<grid, 256>>>(static_cast<const float2*>(source), static_cast<uint32_t*>(target), v,
                                          static_cast<int>(k), range_error, static_cast<int>(ns),
                                          static_cast<int>(nc), static_cast<int>(na), static_cast<int>(nt), rows);
  return launch_status();
}

extern "C" rcabeam_status rcabeam_lowrank_expand_device(const void* coded, const void* basis, size_t k, void* out,
                                                        int out_storage, float* pd, size_t nv, size_t nt) {
  if (!coded || (k && !basis) || (!out && !pd) || !nv || !nt || k > kMaxBasis || nt > kMaxFramesLR)
    return RCABEAM_ERROR_ARGUMENT;
  if (out_storage < RCABEAM_STORAGE_C64 || out_storage > RCABEAM_STORAGE_BF16) return RCABEAM_ERROR_ARGUMENT;
  const long long rows = static_cast<long long>(nv);
  const unsigned grid = static_cast<unsigned>((rows + 3) / 4);
  const float2* c = static_cast<const float2*>(coded);
  const double2* v = static_cast<const double2*>(basis);
  const int ki = static_cast<int>(k), ti = static_cast<int>(nt);
  switch (out_storage) {
    case RCABEAM_STORAGE_C64:
      k_expand<float2><<<grid, 256>>>(c, v, ki, static_cast<float2*>(out), pd, ti, rows);
      break;
    case RCABEAM_STORAGE_FP16:
      k_expand<__half2><<<grid, 256>>>(c, v, ki, static_cast<__half2*>(out), pd, ti, rows);
      break;
    default:
      k_expand<bf16x2><<<grid, 256>>>(c, v, ki, static_cast<bf16x2*>(out), pd, ti, rows);
  }
  return launch_status();
}
