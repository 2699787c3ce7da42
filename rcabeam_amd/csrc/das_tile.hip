// LDS-tiled, frame-batched OPW delay-and-sum for MI355X (gfx950): the main
// path of rcabeam_ensemble_device methods 0 and 5, rcabeam_opw_fp16_device,
// rcabeam_opw_bf16_device and rcabeam_opw_mixed_device (reference
// src/rcabeam/ensemble.cu:341-454 computes the same sums with per-tap global
// gathers).
//
// Work decomposition (one 512-thread workgroup = 8 wave64s, two per CU):
//   * a workgroup owns a tile of 64 consecutive voxels (y fastest: on a
//     Cartesian grid, half a y line at fixed x and z) and every frame, so the
//     power reduction is complete inside the workgroup (deterministic pd);
//   * wave w owns voxels 8w .. 8w+7; the 8 lanes of a voxel own frames
//     lane%8 + 8j, j < FPL (all frames at once, up to 208);
//   * loop order aperture -> channel -> angle.  Per channel the workgroup
//     builds the tap table of its 64 voxels x angles once (receive and
//     transmit terms in float64, combined into row, weight and phasor) and,
//     per angle, the exact row window [min first row, max second row] by
//     wave reductions;
//   * per (channel, angle) the window rows x all frames are staged from the
//     packed (channels, angles, samples, frames) layout into LDS with 16-byte
//     loads (one row holds every frame: contiguous in HBM), then every lane
//     interpolates its voxel's tap for its frames from LDS.  Windows longer
//     than max_rows (40 rows, fewer when LDS is short) are staged in chunks
//     that overlap by one row.
//
// Arithmetic per tap and frame:
//   kCoefFloat  : v = a + w (b - a); acc += v * e^{i phi}        (8 FMA)
//   kCoefSplit  : acc += a (P0hi + P0lo) + b (P1hi + P1lo)       (8 dot2, 16-bit samples)
//   kCoefSingle : acc += a P0 + b P1, 16-bit coefficients          (4 dot2)
// with P0 = (1 - w) e^{i phi}, P1 = w e^{i phi}; v_dot2_f32_{f16,bf16}
// multiplies two 16-bit pairs exactly and accumulates in float32.
#include "ensemble.hpp"
#include "rcabeam_mi355x.h"

#include <algorithm>

namespace rcb {
namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / kWave;
constexpr int kLanesPerVoxel = 8;
constexpr int kVoxelsPerWave = kWave / kLanesPerVoxel;
constexpr int kTile = kWaves * kVoxelsPerWave;  // 64 voxels
constexpr int kGroup = kThreads / kTile;        // channels per receive-term group
constexpr int kMaxAngles = 16;
constexpr int kMaxRows = 40;  // rows of one staged window chunk (fewer when LDS is short)
constexpr int kMinRows = 4;
constexpr int kMaxFramesPerLane = 26;
// Dynamic LDS per workgroup: two workgroups share a CU's 160 KiB.
constexpr int kLdsBytes = 80 * 1024;
constexpr int kBig = 0x7fffffff;

typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf162v __attribute__((ext_vector_type(2)));

template <typename T>
struct Word;  // one complex sample as stored in LDS
template <>
struct Word<float2> {
  using type = float2;
};
template <>
struct Word<__half2> {
  using type = uint32_t;
};
template <>
struct Word<bf16x2> {
  using type = uint32_t;
};

template <typename T>
__device__ __forceinline__ float2 widen(typename Word<T>::type w);
template <>
__device__ __forceinline__ float2 widen<float2>(float2 w) {
  return w;
}
template <>
__device__ __forceinline__ float2 widen<__half2>(uint32_t w) {
  const half2v h = __builtin_bit_cast(half2v, w);
  return make_float2(static_cast<float>(h.x), static_cast<float>(h.y));
}
template <>
__device__ __forceinline__ float2 widen<bf16x2>(uint32_t w) {
  return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}

template <typename T>
__device__ __forceinline__ uint32_t pack2(float lo, float hi);
template <>
__device__ __forceinline__ uint32_t pack2<__half2>(float lo, float hi) {
  half2v h;
  h.x = static_cast<_Float16>(lo);
  h.y = static_cast<_Float16>(hi);
  return __builtin_bit_cast(uint32_t, h);
}
template <>
__device__ __forceinline__ uint32_t pack2<bf16x2>(float lo, float hi) {
  return bf16_bits_rne(lo) | (bf16_bits_rne(hi) << 16);
}

template <typename T>
__device__ __forceinline__ float dot2(uint32_t a, uint32_t b, float c);
template <>
__device__ __forceinline__ float dot2<__half2>(uint32_t a, uint32_t b, float c) {
  return __builtin_amdgcn_fdot2(__builtin_bit_cast(half2v, a), __builtin_bit_cast(half2v, b), c, false);
}
template <>
__device__ __forceinline__ float dot2<bf16x2>(uint32_t a, uint32_t b, float c) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf162v, a), __builtin_bit_cast(bf162v, b), c, false);
}

// Tap coefficients of one lane for one (channel, angle).
struct Coef {
  float w, c, s;                  // kCoefFloat
  uint32_t k0r, k0i, k1r, k1i;    // 16-bit pairs: (P.re, -P.im) and (P.im, P.re) of P0, P1
  uint32_t l0r, l0i, l1r, l1i;    // low parts for kCoefSplit
};

template <typename T, int CP>
__device__ __forceinline__ Coef make_coef(const Tap& t) {
  Coef k{};
  k.w = t.w;
  k.c = t.c;
  k.s = t.s;
  if constexpr (CP != kCoefFloat) {
    const float u = 1.0f - t.w;
    const float p0r = u * t.c, p0i = u * t.s, p1r = t.w * t.c, p1i = t.w * t.s;
    k.k0r = pack2<T>(p0r, -p0i);
    k.k0i = pack2<T>(p0i, p0r);
    k.k1r = pack2<T>(p1r, -p1i);
    k.k1i = pack2<T>(p1i, p1r);
    if constexpr (CP == kCoefSplit) {
      const float2 h0 = widen<T>(k.k0r), h1 = widen<T>(k.k1r);
      const float q0r = p0r - h0.x, q0i = p0i + h0.y, q1r = p1r - h1.x, q1i = p1i + h1.y;
      k.l0r = pack2<T>(q0r, -q0i);
      k.l0i = pack2<T>(q0i, q0r);
      k.l1r = pack2<T>(q1r, -q1i);
      k.l1i = pack2<T>(q1i, q1r);
    }
  }
  return k;
}

// acc[j] += tap(frame fl + 8 j) for one frame block; pa / pb point at this
// lane's first frame in the two sample rows of the tap.
template <typename T, int CP, int FPL>
__device__ __forceinline__ void accumulate(const typename Word<T>::type* pa, const typename Word<T>::type* pb,
                                           const Coef& k, float2 (&acc)[FPL]) {
#pragma unroll
  for (int j = 0; j < FPL; ++j) {
    if constexpr (CP == kCoefFloat) {
      const float2 a = widen<T>(pa[kLanesPerVoxel * j]);
      const float2 b = widen<T>(pb[kLanesPerVoxel * j]);
      const float vx = fmaf(k.w, b.x - a.x, a.x);
      const float vy = fmaf(k.w, b.y - a.y, a.y);
      acc[j].x = fmaf(k.c, vx, fmaf(-k.s, vy, acc[j].x));
      acc[j].y = fmaf(k.c, vy, fmaf(k.s, vx, acc[j].y));
    } else {
      const uint32_t a = pa[kLanesPerVoxel * j];
      const uint32_t b = pb[kLanesPerVoxel * j];
      float x = acc[j].x, y = acc[j].y;
      if constexpr (CP == kCoefSplit) {
        x = dot2<T>(a, k.l0r, x);
        y = dot2<T>(a, k.l0i, y);
        x = dot2<T>(b, k.l1r, x);
        y = dot2<T>(b, k.l1i, y);
      }
      x = dot2<T>(a, k.k0r, x);
      y = dot2<T>(a, k.k0i, y);
      x = dot2<T>(b, k.k1r, x);
      y = dot2<T>(b, k.k1i, y);
      acc[j] = make_float2(x, y);
    }
  }
}

__device__ __forceinline__ int wave_min(int x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = min(x, __shfl_xor(x, off, kWave));
  return x;
}
__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = max(x, __shfl_xor(x, off, kWave));
  return x;
}

struct TileLayout {  // byte offsets inside the dynamic LDS allocation
  int trig, bounds, win, bytes;
};

// taps: Tap[na][kTile] at offset 0; trig: double[na][2]; bounds: int2[na];
// win: W[max_rows][stride].
__host__ __device__ inline TileLayout tile_layout(int na, int stride, int word_bytes, int max_rows) {
  TileLayout l;
  l.trig = na * kTile * static_cast<int>(sizeof(Tap));
  l.bounds = l.trig + kMaxAngles * 2 * static_cast<int>(sizeof(double));
  l.win = (l.bounds + kMaxAngles * 2 * 4 + 15) & ~15;
  l.bytes = l.win + max_rows * stride * word_bytes;
  return l;
}

// Stage rows [r0, r0 + rows) x frames [0, nt) of one (channel, angle) plane
// into win (frames nt .. FB - 1 of the window stay zero from the kernel's
// initial fill).  vec: 16-byte units (the plane and every row start are
// 16-byte aligned, nt is a multiple of the unit).
template <typename W, int FB>
__device__ __forceinline__ void stage(const W* __restrict__ plane, int nt, int r0, int rows, W* win, int stride,
                                      bool vec) {
  const W* __restrict__ src = plane + static_cast<size_t>(r0) * nt;
  if (vec) {
    constexpr int U = 16 / static_cast<int>(sizeof(W));
    constexpr int UPR = FB / U;  // units per row
    const int n = rows * UPR;
    for (int q = threadIdx.x; q < n; q += kThreads) {
      const int row = q / UPR;
      const int col = (q - row * UPR) * U;
      if (col < nt)
        *reinterpret_cast<uint4*>(win + row * stride + col) =
            *reinterpret_cast<const uint4*>(src + row * nt + col);
    }
  } else {
    const int n = rows * FB;
    for (int q = threadIdx.x; q < n; q += kThreads) {
      const int row = q / FB;
      const int col = q - row * FB;
      if (col < nt) win[row * stride + col] = src[row * nt + col];
    }
  }
}

template <typename O>
__device__ __forceinline__ void store_out(void* signal, long long index, float2 v) {
  store_sample(reinterpret_cast<O*>(signal) + index, v);
}

template <typename T, int CP, int FPL, bool APERTURES>
__global__ __launch_bounds__(kThreads) void das_opw_tile(EnsembleArgs p, long long tiles, int stride, int max_rows,
                                                         bool vec, int out_storage) {
  using W = typename Word<T>::type;
  constexpr int FB = kLanesPerVoxel * FPL;
  extern __shared__ __align__(16) unsigned char lds[];
  const TileLayout layout = tile_layout(p.na, stride, static_cast<int>(sizeof(W)), max_rows);
  Tap* taps = reinterpret_cast<Tap*>(lds);                        // [na][kTile]
  double* trig = reinterpret_cast<double*>(lds + layout.trig);    // [na][cos, sin]
  int2* bounds = reinterpret_cast<int2*>(lds + layout.bounds);    // [na]: first row, last row
  W* win = reinterpret_cast<W*>(lds + layout.win);                // [max_rows][stride]

  const long long tile = xcd_tile(tiles);
  if (tile >= tiles) return;
  const long long v0 = tile * kTile;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int vl = wave * kVoxelsPerWave + lane / kLanesPerVoxel;  // voxel of this lane's taps
  const int fl = lane % kLanesPerVoxel;                          // frame lane
  const long long v = v0 + vl;

  // Voxel v0 + lane: this thread builds its taps for angles wave, wave + 8.
  const long long vr = v0 + lane;
  const bool vr_ok = vr < p.nv;
  double xr = 0.0, zr = 0.0, yr = 0.0;
  if (vr_ok) {
    xr = p.scan[3 * vr];
    zr = p.scan[3 * vr + 1];
    yr = p.scan[3 * vr + 2];
  }
  // Steering cosines and sines once per workgroup (the float64 sincos is
  // register-hungry, so it runs before the accumulators are live); the
  // window is zeroed once, so frames nt .. FB - 1 read as zero.
  if (tid < p.na) {
    double sn, cs;
    sincos(static_cast<double>(p.angles[tid]), &sn, &cs);
    trig[2 * tid] = cs;
    trig[2 * tid + 1] = sn;
  }
  {
    uint4* w4 = reinterpret_cast<uint4*>(win);
    const int n = max_rows * stride * static_cast<int>(sizeof(W)) / 16;
    for (int i = tid; i < n; i += kThreads) w4[i] = make_uint4(0u, 0u, 0u, 0u);
  }

  float2 acc[FPL];
#pragma unroll
  for (int j = 0; j < FPL; ++j) acc[j] = make_float2(0.f, 0.f);

  for (int cfg = 0; cfg < 2; ++cfg) {
    const W* __restrict__ data = static_cast<const W*>(cfg == 0 ? p.rc : p.cr);
    const float* el = cfg == 0 ? p.xe : p.ye;
    for (int ch = 0; ch < p.nc; ++ch) {
      __syncthreads();  // the previous channel's angles are done with taps, bounds and window
      {
        // Tap table of this channel and the exact row window of every angle.
        Term r{kInactive, 0.f, 0.f, 0.f};
        if (vr_ok) r = receive_term(cfg == 0 ? xr : yr, zr, static_cast<double>(el[ch]), p.c, p.fs, p.fd, p.fn);
        for (int a = wave; a < p.na; a += kWaves) {
          Tap tp{-1, 0.f, 0.f, 0.f};
          if (r.i != kInactive) {
            const Term t = transmit_term(cfg == 0 ? yr : xr, zr, trig[2 * a], trig[2 * a + 1],
                                         static_cast<double>(p.starts[a]), p.c, p.fs, p.fd);
            tp = combine(r, t, p.ns);
          }
          taps[a * kTile + lane] = tp;
          const bool live = tp.i0 >= 0;
          const int lo = wave_min(live ? tp.i0 : kBig);
          const int hi = wave_max(live ? tp.i0 + (tp.w != 0.f ? 1 : 0) : -kBig);
          if (lane == 0) bounds[a] = make_int2(lo, hi);
        }
      }
      __syncthreads();
      bool staged = false;  // a stage since the last barrier: the next one must wait for its readers
      for (int a = 0; a < p.na; ++a) {
        const int2 b = bounds[a];
        const int lo = uniform(b.x);
        const int hi = uniform(b.y);
        if (lo > hi) continue;  // no tap of the tile for this (channel, angle)
        const Tap tp = taps[a * kTile + vl];
        const bool live = tp.i0 >= 0;
        const int i1 = tp.w != 0.f ? tp.i0 + 1 : tp.i0;
        const Coef k = make_coef<T, CP>(tp);
        const W* plane = data + static_cast<size_t>(ch * p.na + a) * p.ns * p.nt;
        for (int r0 = lo;; r0 += max_rows - 1) {
          const bool last = r0 + max_rows - 1 >= hi;
          const int rows = last ? hi + 1 - r0 : max_rows;
          const bool in_chunk = live && tp.i0 >= r0 && (last || tp.i0 < r0 + max_rows - 1);
          if (staged) __syncthreads();
          stage<W, FB>(plane, p.nt, r0, rows, win, stride, vec);
          __syncthreads();
          staged = true;
          if (in_chunk)
            accumulate<T, CP, FPL>(win + (tp.i0 - r0) * stride + fl, win + (i1 - r0) * stride + fl, k, acc);
          if (last) break;
        }
      }
    }
    if constexpr (APERTURES) {
      if (v < p.nv) {
#pragma unroll
        for (int j = 0; j < FPL; ++j) {
          const int t = fl + kLanesPerVoxel * j;
          if (t < p.nt) p.signal[(2 * v + cfg) * p.nt + t] = acc[j];
          acc[j] = make_float2(0.f, 0.f);
        }
        if (fl == 0) {
          p.pd[2 * v] = 0.f;
          p.pd[2 * v + 1] = 0.f;
        }
      }
    }
  }
  if constexpr (!APERTURES) {
    float power = 0.f;
#pragma unroll
    for (int j = 0; j < FPL; ++j) {
      const int t = fl + kLanesPerVoxel * j;
      if (t < p.nt && v < p.nv) {
        const float2 s = acc[j];
        power = fmaf(s.x, s.x, fmaf(s.y, s.y, power));
        if (p.signal) {
          const long long index = v * p.nt + t;
          if (out_storage == RCABEAM_STORAGE_FP16)
            store_out<__half2>(p.signal, index, s);
          else if (out_storage == RCABEAM_STORAGE_BF16)
            store_out<bf16x2>(p.signal, index, s);
          else
            store_out<float2>(p.signal, index, s);
        }
      }
    }
    // The 8 lanes of a voxel are adjacent: butterfly inside the group.
#pragma unroll
    for (int off = 1; off < kLanesPerVoxel; off <<= 1) power += __shfl_xor(power, off, kWave);
    if (fl == 0 && v < p.nv) {
      p.pd[2 * v] = power / static_cast<float>(p.nt);
      p.pd[2 * v + 1] = 0.f;
    }
  }
}

// Row stride of the LDS window (in samples): at least FB and such that the 8
// rows a wave reads in one instruction fall on distinct LDS bank groups.
int window_stride(int fb, int word_bytes) {
  const int seg = kLanesPerVoxel * word_bytes / 4;  // dwords one voxel's lanes read
  int dw = (fb * word_bytes / 4 + seg - 1) / seg * seg;
  if ((dw / seg) % 2 == 0) dw += seg;
  return dw * 4 / word_bytes;
}

template <typename T, int CP, int FPL, bool APERTURES>
hipError_t launch_shape(const EnsembleArgs& p, int out_storage) {
  using W = typename Word<T>::type;
  constexpr int FB = kLanesPerVoxel * FPL;
  const int stride = window_stride(FB, static_cast<int>(sizeof(W)));
  const int fixed = tile_layout(p.na, stride, static_cast<int>(sizeof(W)), 0).bytes;
  const int max_rows = std::min(kMaxRows, (kLdsBytes - fixed) / (stride * static_cast<int>(sizeof(W))));
  if (max_rows < kMinRows) return hipErrorNotSupported;
  const TileLayout layout = tile_layout(p.na, stride, static_cast<int>(sizeof(W)), max_rows);
  const long long tiles = (p.nv + kTile - 1) / kTile;
  const long long grid = (tiles + 7) / 8 * 8;
  if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
  constexpr int U = 16 / static_cast<int>(sizeof(W));
  const bool vec = reinterpret_cast<uintptr_t>(p.rc) % 16 == 0 && reinterpret_cast<uintptr_t>(p.cr) % 16 == 0 &&
                   p.nt % U == 0;
  auto kernel = das_opw_tile<T, CP, FPL, APERTURES>;
  static bool attribute_set = false;  // one opt-in per instantiation (beyond 64 KiB of dynamic LDS)
  if (!attribute_set) {
    const hipError_t e =
        hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                            kLdsBytes);
    if (e != hipSuccess) return e;
    attribute_set = true;
  }
  kernel<<<dim3(static_cast<unsigned>(grid)), kThreads, layout.bytes>>>(p, tiles, stride, max_rows, vec,
                                                                        out_storage);
  return hipGetLastError();
}

template <typename T, int CP, bool APERTURES>
hipError_t launch_frames(const EnsembleArgs& p, int out_storage) {
  if (p.nt <= 16) return launch_shape<T, CP, 2, APERTURES>(p, out_storage);
  if (p.nt <= 32) return launch_shape<T, CP, 4, APERTURES>(p, out_storage);
  if (p.nt <= 64) return launch_shape<T, CP, 8, APERTURES>(p, out_storage);
  if (p.nt <= 104) return launch_shape<T, CP, 13, APERTURES>(p, out_storage);
  return launch_shape<T, CP, 26, APERTURES>(p, out_storage);
}

}  // namespace

hipError_t launch_opw_tile(const EnsembleArgs& p, int channel_storage, int coefficients, int signal_storage,
                           bool apertures) {
  if (p.na > kMaxAngles || p.nt > kLanesPerVoxel * kMaxFramesPerLane) return hipErrorNotSupported;
  // Plane offsets and in-plane row offsets are int.
  if (static_cast<long long>(p.ns) * p.nt > 0x7fffffffLL) return hipErrorNotSupported;
  if (signal_storage < RCABEAM_STORAGE_C64 || signal_storage > RCABEAM_STORAGE_BF16) return hipErrorInvalidValue;
  if (apertures) {
    if (channel_storage != RCABEAM_STORAGE_C64) return hipErrorInvalidValue;
    return launch_frames<float2, kCoefFloat, true>(p, signal_storage);
  }
  switch (channel_storage) {
    case RCABEAM_STORAGE_C64:
      return launch_frames<float2, kCoefFloat, false>(p, signal_storage);
    case RCABEAM_STORAGE_FP16:
      if (coefficients == kCoefSingle) return launch_frames<__half2, kCoefSingle, false>(p, signal_storage);
      if (coefficients == kCoefSplit) return launch_frames<__half2, kCoefSplit, false>(p, signal_storage);
      return launch_frames<__half2, kCoefFloat, false>(p, signal_storage);
    case RCABEAM_STORAGE_BF16:
      if (coefficients == kCoefSingle) return launch_frames<bf16x2, kCoefSingle, false>(p, signal_storage);
      return launch_frames<bf16x2, kCoefFloat, false>(p, signal_storage);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace rcb
