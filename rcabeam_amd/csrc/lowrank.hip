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
#include <algorithm>

#include "common.hpp"
#include "rcabeam_mi355x.h"
#include "tile_staging.hpp"

namespace rcb {
namespace {

constexpr int kMaxBasis = 12;        // k <= 12 keeps nt + 2k <= 224 for nt = 200
constexpr int kMaxFramesLR = 256;    // 32 frames per lane, 8 lanes per row

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

// Sixteen lanes per target (channel, angle, sample) row: lane f holds frames
// f + 16 j (j < FPL), so one load instruction reads 128 contiguous bytes of a
// row and the K coordinate reductions stay inside 16 lanes; 16 rows per
// 256-thread workgroup, the basis staged in LDS.  K and FPL are compile-time
// bounds of k and of the frames per lane (registers for those alone: 191
// VGPRs when they were sized for k <= 12 and 256 frames).  Coordinates and
// residual in float64 (the row is read once; HBM-bound).
constexpr int kPackLanes = 16;
constexpr int kPackRowsPerBlock = 256 / kPackLanes;
// The expansion: 16 lanes per row, 16 rows per 256-thread workgroup.
constexpr int kExpandLanes = 16;
constexpr int kExpandRowsPerBlock = 256 / kExpandLanes;

template <typename T, int K, int FPL>
__global__ __launch_bounds__(256) void k_pack_lowrank(const float2* __restrict__ src, uint32_t* __restrict__ dst,
                                                      const double2* __restrict__ basis, int k, int* range_error,
                                                      int ns, int nc, int na, int nt, long long rows) {
  extern __shared__ double2 vs[];  // [k][nt]
  for (int i = threadIdx.x; i < k * nt; i += blockDim.x) vs[i] = basis[i];
  __syncthreads();
  const long long row = static_cast<long long>(blockIdx.x) * kPackRowsPerBlock + threadIdx.x / kPackLanes;
  const int f = threadIdx.x & (kPackLanes - 1);
  bool bad = false;
  if (row < rows) {
    const long long s = row % ns;
    const long long ca = row / ns;
    const long long a = ca % na;
    const long long ch = ca / na;
    const float2* in = src + ((s * nc + ch) * na + a) * nt;
    const int ntc = nt + 2 * k;
    uint32_t* out = dst + row * ntc;
    float2 x[FPL];
#pragma unroll
    for (int j = 0; j < FPL; ++j) {
      const int t = f + kPackLanes * j;
      x[j] = t < nt ? in[t] : make_float2(0.f, 0.f);
    }
    double wr[K], wi[K];  // w_q = sum_t c_t V[t, q], in every lane of the row
#pragma unroll
    for (int q = 0; q < K; ++q) {
      wr[q] = wi[q] = 0.0;
      if (q < k) {
#pragma unroll
        for (int j = 0; j < FPL; ++j) {
          const int t = f + kPackLanes * j;
          if (t < nt) {
            const double2 b = vs[q * nt + t];
            wr[q] += static_cast<double>(x[j].x) * b.x - static_cast<double>(x[j].y) * b.y;
            wi[q] += static_cast<double>(x[j].x) * b.y + static_cast<double>(x[j].y) * b.x;
          }
        }
#pragma unroll
        for (int off = 1; off < kPackLanes; off <<= 1) {
          wr[q] += __shfl_xor(wr[q], off, kWave);
          wi[q] += __shfl_xor(wi[q], off, kWave);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < FPL; ++j) {
      const int t = f + kPackLanes * j;
      if (t < nt) {  // e_t = c_t - sum_q w_q conj(V[t, q])
        double er = x[j].x, ei = x[j].y;
#pragma unroll
        for (int q = 0; q < K; ++q) {
          if (q < k) {
            const double2 b = vs[q * nt + t];
            er -= wr[q] * b.x + wi[q] * b.y;
            ei -= wi[q] * b.x - wr[q] * b.y;
          }
        }
        const float re = static_cast<float>(er), im = static_cast<float>(ei);
        bad |= !in_range<T>(re) || !in_range<T>(im);
        out[t] = pack2<T>(re, im);
      }
    }
    // Frames nt + 2q (hi) and nt + 2q + 1 (lo) hold w_q; lane f writes q = f.
#pragma unroll
    for (int q = 0; q < K; ++q) {
      if (q < k && q == f) {
        const float re = static_cast<float>(wr[q]), im = static_cast<float>(wi[q]);
        bad |= !in_range<T>(re) || !in_range<T>(im);
        const uint32_t hi = pack2<T>(re, im);
        const float2 h = widen<T>(hi);
        out[nt + 2 * q] = hi;
        out[nt + 2 * q + 1] = pack2<T>(static_cast<float>(wr[q] - h.x), static_cast<float>(wi[q] - h.y));
      }
    }
  }
  if (range_error && __any(bad) && (threadIdx.x & (kWave - 1)) == 0) *range_error = 1;
}

template <typename T, int K>
void launch_pack_lowrank(unsigned grid, size_t lds, const float2* src, uint32_t* dst, const double2* v, int k,
                         int* range_error, int ns, int nc, int na, int nt, long long rows) {
  if (nt <= 13 * kPackLanes)
    k_pack_lowrank<T, K, 13><<<grid, 256, lds>>>(src, dst, v, k, range_error, ns, nc, na, nt, rows);
  else
    k_pack_lowrank<T, K, kMaxFramesLR / kPackLanes><<<grid, 256, lds>>>(src, dst, v, k, range_error, ns, nc, na,
                                                                         nt, rows);
}

template <typename T>
void launch_pack_lowrank_k(unsigned grid, size_t lds, const float2* src, uint32_t* dst, const double2* v, int k,
                           int* range_error, int ns, int nc, int na, int nt, long long rows) {
  if (k <= 4)
    launch_pack_lowrank<T, 4>(grid, lds, src, dst, v, k, range_error, ns, nc, na, nt, rows);
  else if (k <= 8)
    launch_pack_lowrank<T, 8>(grid, lds, src, dst, v, k, range_error, ns, nc, na, nt, rows);
  else
    launch_pack_lowrank<T, kMaxBasis>(grid, lds, src, dst, v, k, range_error, ns, nc, na, nt, rows);
}

// Sixteen lanes per voxel row of the coded delay-and-sum output (nt + 2k
// complex64 columns), lane f holding frames f + 16 j, so one load reads 128
// contiguous bytes of a row; 16 rows per workgroup, the basis staged in LDS
// as float32 (converted once per workgroup, the values each frame used to
// convert): x_t = d_t + sum_j (d[nt + 2j] + d[nt + 2j + 1]) conj(V[t, j]),
// stored in `O`; power (mean over the nt frames of |x_t|^2) into pd[2v].
template <typename O>
__global__ __launch_bounds__(256) void k_expand(const float2* __restrict__ coded, const double2* __restrict__ basis,
                                                int k, O* __restrict__ out, float* __restrict__ pd, int nt,
                                                long long rows) {
  extern __shared__ float2 vf[];  // [k][nt]
  for (int i = threadIdx.x; i < k * nt; i += blockDim.x) {
    const double2 b = basis[i];
    vf[i] = make_float2(static_cast<float>(b.x), static_cast<float>(b.y));
  }
  __syncthreads();
  const long long row = static_cast<long long>(blockIdx.x) * kExpandRowsPerBlock + threadIdx.x / kExpandLanes;
  const int f = threadIdx.x & (kExpandLanes - 1);
  if (row >= rows) return;  // after the only barrier
  const int ntc = nt + 2 * k;
  const float2* in = coded + row * ntc;
  // The delay-and-sum of w_j (hi + lo parts), read by lane j and shared in the row's 16 lanes.
  float dwr[kMaxBasis], dwi[kMaxBasis];
#pragma unroll
  for (int q = 0; q < kMaxBasis; ++q) {
    float2 d = make_float2(0.f, 0.f);
    if (q < k && q == f) {
      const float2 h = in[nt + 2 * q], l = in[nt + 2 * q + 1];
      d = make_float2(h.x + l.x, h.y + l.y);
    }
    dwr[q] = __shfl(d.x, (threadIdx.x & (kWave - kExpandLanes)) | q, kWave);
    dwi[q] = __shfl(d.y, (threadIdx.x & (kWave - kExpandLanes)) | q, kWave);
  }
  float power = 0.f;
#pragma unroll 4
  for (int j = 0; j < kMaxFramesLR / kExpandLanes; ++j) {
    const int t = f + kExpandLanes * j;
    if (t < nt) {
      float2 x = in[t];
#pragma unroll
      for (int q = 0; q < kMaxBasis; ++q) {
        if (q < k) {
          const float2 b = vf[q * nt + t];
          x.x = fmaf(dwr[q], b.x, fmaf(dwi[q], b.y, x.x));  // + w conj(b)
          x.y = fmaf(dwi[q], b.x, fmaf(-dwr[q], b.y, x.y));
        }
      }
      power = fmaf(x.x, x.x, fmaf(x.y, x.y, power));
      if (out) store_sample(out + row * nt + t, x);
    }
  }
#pragma unroll
  for (int off = 1; off < kExpandLanes; off <<= 1) power += __shfl_xor(power, off, kWave);
  if (pd && f == 0) {
    pd[2 * row] = power / static_cast<float>(nt);
    pd[2 * row + 1] = 0.f;
  }
}

// Dominant k-dimensional invariant subspace of a Hermitian nt x nt complex128
// matrix G (column-major) by subspace iteration: from k fixed pseudo-random
// vectors, `iterations` times W = G V followed by modified Gram-Schmidt (twice
// on the last pass), in one workgroup with V and W in LDS.  A zero G gives a
// zero basis.  The result is an
// orthonormal basis whose span approaches that of the k leading eigenvectors
// at the rate (lambda_{k+1} / lambda_k)^iterations; the low-rank coding needs
// only a basis that catches the strong clutter modes (any orthonormal basis
// codes the channels exactly).
constexpr int kSubThreads = 1024;
constexpr int kSubWaves = kSubThreads / kWave;
constexpr int kSubMaxK = 12;
constexpr int kSubMaxElems = 4096;  // k * nt per LDS array (64 KiB of complex128)

__device__ __forceinline__ double2 cmulc_d(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// Sum over the workgroup of n <= kSubMaxK complex values per thread; every
// thread receives the sums.  Ends with a barrier.
__device__ void block_sum(double2 (&v)[kSubMaxK], int n, double2* red) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  for (int j = 0; j < n; ++j) {
    double x = v[j].x, y = v[j].y;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      x += __shfl_xor(x, off, kWave);
      y += __shfl_xor(y, off, kWave);
    }
    if (lane == 0) red[wave * kSubMaxK + j] = make_double2(x, y);
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    double x = 0.0, y = 0.0;
    for (int w = 0; w < kSubWaves; ++w) {  // fixed order: the same sums on every thread
      x += red[w * kSubMaxK + j].x;
      y += red[w * kSubMaxK + j].y;
    }
    v[j] = make_double2(x, y);
  }
  __syncthreads();
}

// Orthonormalize the k columns of w (rows of [k][nt]) into v by modified
// Gram-Schmidt; w is overwritten.  A zero column stays zero.
__device__ void mgs(double2* w, double2* v, int nt, int k, double2* red) {
  for (int a = 0; a < k; ++a) {
    double2 part[kSubMaxK];
    part[0] = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < nt; i += kSubThreads) {
      const double2 x = w[a * nt + i];
      part[0].x += x.x * x.x + x.y * x.y;
    }
    block_sum(part, 1, red);
    const double inv = part[0].x > 0.0 ? 1.0 / sqrt(part[0].x) : 0.0;
    for (int i = threadIdx.x; i < nt; i += kSubThreads) {
      const double2 x = w[a * nt + i];
      v[a * nt + i] = make_double2(x.x * inv, x.y * inv);
    }
    __syncthreads();
    const int rest = k - a - 1;
    if (rest == 0) break;
    for (int b = 0; b < rest; ++b) part[b] = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < nt; i += kSubThreads) {
      const double2 va = v[a * nt + i];
      for (int b = 0; b < rest; ++b) {
        const double2 d = cmulc_d(va, w[(a + 1 + b) * nt + i]);
        part[b].x += d.x;
        part[b].y += d.y;
      }
    }
    block_sum(part, rest, red);
    for (int i = threadIdx.x; i < nt; i += kSubThreads) {
      const double2 va = v[a * nt + i];
      for (int b = 0; b < rest; ++b) {
        double2& x = w[(a + 1 + b) * nt + i];
        x.x -= part[b].x * va.x - part[b].y * va.y;
        x.y -= part[b].x * va.y + part[b].y * va.x;
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kSubThreads) void k_subspace(const double2* __restrict__ g, int nt, int k, int iterations,
                                                          double2* __restrict__ basis) {
  extern __shared__ double2 sub_lds[];
  double2* v = sub_lds;           // [k][nt]: the basis, row c = column c of V
  double2* w = v + k * nt;        // [k][nt]
  double2* red = w + k * nt;      // [kSubWaves][kSubMaxK]
  // Start from fixed pseudo-random vectors (entries in [-1, 1) from an integer hash), so that no column of G
  // that happens to be zero or dependent can leave a start vector without a dominant component.
  for (int idx = threadIdx.x; idx < k * nt; idx += kSubThreads) {
    unsigned h = static_cast<unsigned>(idx) * 0x9E3779B9u + 0x7F4A7C15u;
    h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16;
    v[idx] = make_double2(static_cast<double>(h & 0xFFFFu) / 32768.0 - 1.0, static_cast<double>(h >> 16) / 32768.0 - 1.0);
  }
  __syncthreads();
  for (int it = 0; it < iterations; ++it) {
    for (int idx = threadIdx.x; idx < k * nt; idx += kSubThreads) {
      const int c = idx / nt, i = idx - c * nt;
      double re = 0.0, im = 0.0;
      for (int j = 0; j < nt; ++j) {
        const double2 gij = g[i + static_cast<long long>(j) * nt];
        const double2 vj = v[c * nt + j];
        re = fma(gij.x, vj.x, fma(-gij.y, vj.y, re));
        im = fma(gij.x, vj.y, fma(gij.y, vj.x, im));
      }
      w[idx] = make_double2(re, im);
    }
    __syncthreads();
    mgs(w, v, nt, k, red);
  }
  // Second pass on the result: orthonormal to working precision.
  for (int idx = threadIdx.x; idx < k * nt; idx += kSubThreads) w[idx] = v[idx];
  __syncthreads();
  mgs(w, v, nt, k, red);
  for (int idx = threadIdx.x; idx < k * nt; idx += kSubThreads) basis[idx] = v[idx];
}

}  // namespace
}  // namespace rcb

using namespace rcb;

extern "C" rcabeam_status rcabeam_dominant_subspace_device(const void* gram, size_t frames, size_t k,
                                                           size_t iterations, void* basis) {
  if (!gram || !basis || !frames || !k || k > static_cast<size_t>(kSubMaxK) || k >= frames ||
      k * frames > static_cast<size_t>(kSubMaxElems) || !iterations || iterations > 1000 ||
      reinterpret_cast<uintptr_t>(gram) % 16 || reinterpret_cast<uintptr_t>(basis) % 16)
    return RCABEAM_ERROR_ARGUMENT;
  const size_t lds = (2 * k * frames + kSubWaves * kSubMaxK) * sizeof(double2);
  const hipError_t e = allow_dynamic_lds(reinterpret_cast<const void*>(&k_subspace), static_cast<int>(lds));
  if (e != hipSuccess) return RCABEAM_ERROR_CUDA;
  k_subspace<<<1, kSubThreads, lds>>>(static_cast<const double2*>(gram), static_cast<int>(frames), static_cast<int>(k),
                                       static_cast<int>(iterations), static_cast<double2*>(basis));
  return launch_status();
}

extern "C" rcabeam_status rcabeam_pack_ensemble_lowrank_device(const void* source, void* target, int storage,
                                                               const void* basis, size_t k, int* range_error,
                                                               size_t ns, size_t nc, size_t na, size_t nt) {
  if (!source || !target || !range_error || (k && !basis) || !ns || !nc || !na || !nt) return RCABEAM_ERROR_ARGUMENT;
  if (k > kMaxBasis || nt > kMaxFramesLR || ns * nc * na > (1ull << 40)) return RCABEAM_ERROR_ARGUMENT;
  if (storage != RCABEAM_STORAGE_FP16 && storage != RCABEAM_STORAGE_BF16) return RCABEAM_ERROR_ARGUMENT;
  const long long rows = static_cast<long long>(ns * nc * na);
  const long long blocks = (rows + kPackRowsPerBlock - 1) / kPackRowsPerBlock;
  if (blocks > 0x7fffffffLL) return RCABEAM_ERROR_ARGUMENT;
  const unsigned grid = static_cast<unsigned>(blocks);
  const size_t lds = std::max<size_t>(k * nt * sizeof(double2), sizeof(double2));  // <= 48 KiB
  const double2* v = static_cast<const double2*>(basis);
  const float2* src = static_cast<const float2*>(source);
  uint32_t* dst = static_cast<uint32_t*>(target);
  if (storage == RCABEAM_STORAGE_FP16)
    launch_pack_lowrank_k<__half2>(grid, lds, src, dst, v, static_cast<int>(k), range_error, static_cast<int>(ns),
                                   static_cast<int>(nc), static_cast<int>(na), static_cast<int>(nt), rows);
  else
    launch_pack_lowrank_k<bf16x2>(grid, lds, src, dst, v, static_cast<int>(k), range_error, static_cast<int>(ns),
                                  static_cast<int>(nc), static_cast<int>(na), static_cast<int>(nt), rows);
  return launch_status();
}

extern "C" rcabeam_status rcabeam_lowrank_expand_device(const void* coded, const void* basis, size_t k, void* out,
                                                        int out_storage, float* pd, size_t nv, size_t nt) {
  if (!coded || (k && !basis) || (!out && !pd) || !nv || !nt || k > kMaxBasis || nt > kMaxFramesLR)
    return RCABEAM_ERROR_ARGUMENT;
  if (out_storage < RCABEAM_STORAGE_C64 || out_storage > RCABEAM_STORAGE_BF16) return RCABEAM_ERROR_ARGUMENT;
  const long long rows = static_cast<long long>(nv);
  const unsigned grid = static_cast<unsigned>((rows + kExpandRowsPerBlock - 1) / kExpandRowsPerBlock);
  const size_t lds = std::max<size_t>(k * nt * sizeof(float2), sizeof(float2));
  const float2* c = static_cast<const float2*>(coded);
  const double2* v = static_cast<const double2*>(basis);
  const int ki = static_cast<int>(k), ti = static_cast<int>(nt);
  switch (out_storage) {
    case RCABEAM_STORAGE_C64:
      k_expand<float2><<<grid, 256, lds>>>(c, v, ki, static_cast<float2*>(out), pd, ti, rows);
      break;
    case RCABEAM_STORAGE_FP16:
      k_expand<__half2><<<grid, 256, lds>>>(c, v, ki, static_cast<__half2*>(out), pd, ti, rows);
      break;
    default:
      k_expand<bf16x2><<<grid, 256, lds>>>(c, v, ki, static_cast<bf16x2*>(out), pd, ti, rows);
  }
  return launch_status();
}
