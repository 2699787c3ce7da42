// LDS-tiled, frame-batched OPW delay-and-sum for MI355X (gfx950): the main
// path of rcabeam_ensemble_device methods 0 and 5, rcabeam_opw_fp16_device,
// rcabeam_opw_bf16_device, rcabeam_opw_mixed_device and
// rcabeam_opw_grid_device (reference src/rcabeam/ensemble.cu:341-454 computes
// the same sums with per-tap global gathers).
//
// Work decomposition (one 512-thread workgroup = 8 wave64s, two per CU):
//   * a workgroup owns a tile of 64 voxels and every frame, so the power
//     reduction is complete inside the workgroup (deterministic pd).  On a
//     Cartesian grid a tile is an 8 (x) by 8 (y) patch at one depth, tiles
//     ordered depth-fastest; for a voxel list it is 64 consecutive voxels;
//   * wave w owns voxels 8w .. 8w+7 (one x row of a grid tile); the 8 lanes
//     of a voxel own frames lane%8 + 8j, j < FPL (all frames, up to 208);
//   * loop order aperture -> channel -> angle.  Per channel the workgroup
//     builds the tap table of its 64 voxels x angles once (receive and
//     transmit terms in float64, combined into row, weight and phasor) and,
//     per angle, the exact row window [min first row, max second row] by
//     wave reductions; the next channel's table is built while the current
//     one streams.  Receive rows move by < 1.2 samples per voxel in x and
//     transmit rows by < 0.5 in y, so a patch's window is a few rows;
//   * the windows of consecutive angles are packed into stages of at most
//     max_rows rows (longer windows are split into chunks overlapping by one
//     row).  A stage streams rows x frames from the packed (channels, angles,
//     samples, frames) layout straight into LDS (global_load_lds_dwordx4, no
//     register staging), double-buffered across angles and channels: the
//     next stage is in flight while every lane interpolates its voxel's taps
//     for its frames from the current one; one barrier per stage.
//
// Arithmetic per tap and frame:
//   kCoefFloat  : v = a + w (b - a); acc += v * e^{i phi}        (8 FMA)
//   kCoefSplit  : acc += a (P0hi + P0lo) + b (P1hi + P1lo)       (8 dot2, 16-bit samples)
//   kCoefSingle : acc += a P0 + b P1, 16-bit coefficients          (4 dot2)
// with P0 = (1 - w) e^{i phi}, P1 = w e^{i phi}; v_dot2_f32_{f16,bf16}
// multiplies two 16-bit pairs exactly and accumulates in float32.
#include "ensemble.hpp"
#include "rcabeam_mi355x.h"
#include "tile_staging.hpp"

#include <algorithm>

namespace rcb {
namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / kWave;
constexpr int kLanesPerVoxel = 8;
constexpr int kVoxelsPerWave = kWave / kLanesPerVoxel;
constexpr int kTile = kWaves * kVoxelsPerWave;  // 64 voxels
constexpr int kPatch = 8;                       // grid tiles: kPatch x kPatch lateral voxels
constexpr int kMaxAngles = 16;
constexpr int kMaxRows = 40;  // rows of one stage (fewer when LDS is short)
constexpr int kMinRows = 4;
constexpr int kMaxFramesPerLane = 26;
// Dynamic LDS per workgroup: two workgroups share a CU's 160 KiB.
constexpr int kLdsBytes = 80 * 1024;

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

// Row stride of an LDS window (in samples): at least fb and such that the 8
// rows a wave reads in one instruction fall on distinct LDS bank groups.
constexpr int window_stride(int fb, int word_bytes) {
  const int seg = kLanesPerVoxel * word_bytes / 4;  // dwords one voxel's lanes read
  const int dw0 = (fb * word_bytes / 4 + seg - 1) / seg * seg;
  const int dw = (dw0 / seg) % 2 == 0 ? dw0 + seg : dw0;
  return dw * 4 / word_bytes;
}

struct TileLayout {  // byte offsets inside the dynamic LDS allocation
  int trig, bounds, win, bytes;
};

// taps: Tap[2][na][kTile] at offset 0 (channels alternate); trig:
// double[kMaxAngles][2]; bounds: int2[2][kMaxAngles]; win: two buffers
// W[max_rows][stride].
__host__ __device__ inline TileLayout tile_layout(int na, int stride, int word_bytes, int max_rows) {
  TileLayout l;
  l.trig = 2 * na * kTile * static_cast<int>(sizeof(Tap));
  l.bounds = l.trig + kMaxAngles * 2 * static_cast<int>(sizeof(double));
  l.win = (l.bounds + 2 * kMaxAngles * 2 * 4 + 15) & ~15;
  l.bytes = l.win + 2 * max_rows * stride * word_bytes;
  return l;
}

// Which voxels a tile holds.  grid: 8 x 8 lateral patches, one depth per tile,
// tile = (bx * nby + by) * nz + iz; voxel id (ix * nz + iz) * ny + iy.
// Otherwise voxels [64 t, 64 t + 64) of the scan list.
struct TileMap {
  const float* x;
  const float* z;
  const float* y;
  int nx, nz, ny, nby;
  bool grid;
};

// Voxel id of tile slot s (-1 outside the volume) and its float32 coordinates.
__device__ __forceinline__ long long slot_voxel(const EnsembleArgs& p, const TileMap& m, long long tile, int s,
                                                float& x, float& z, float& y) {
  x = z = y = 0.f;
  if (m.grid) {
    const int iz = static_cast<int>(tile % m.nz);
    const long long lateral = tile / m.nz;
    const int ix = static_cast<int>(lateral / m.nby) * kPatch + s / kPatch;
    const int iy = static_cast<int>(lateral % m.nby) * kPatch + s % kPatch;
    if (ix >= m.nx || iy >= m.ny) return -1;
    x = m.x[ix];
    z = m.z[iz];
    y = m.y[iy];
    return (static_cast<long long>(ix) * m.nz + iz) * m.ny + iy;
  }
  const long long v = tile * kTile + s;
  if (v >= p.nv) return -1;
  x = p.scan[3 * v];
  z = p.scan[3 * v + 1];
  y = p.scan[3 * v + 2];
  return v;
}


template <typename O>
__device__ __forceinline__ void store_out(void* signal, long long index, float2 v) {
  store_sample(reinterpret_cast<O*>(signal) + index, v);
}

// Two workgroups (16 waves) per CU: at most 128 VGPRs.
template <typename T, int CP, int FPL, bool APERTURES>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void das_opw_tile(
    EnsembleArgs p, TileMap map, long long tiles, int max_rows, bool vec, int out_storage) {
  using W = typename Word<T>::type;
  constexpr int FB = kLanesPerVoxel * FPL;
  constexpr int STRIDE = window_stride(FB, static_cast<int>(sizeof(W)));
  extern __shared__ __align__(16) unsigned char lds[];
  const TileLayout layout = tile_layout(p.na, STRIDE, static_cast<int>(sizeof(W)), max_rows);
  Tap* taps = reinterpret_cast<Tap*>(lds);                      // [2][na][kTile]
  double* trig = reinterpret_cast<double*>(lds + layout.trig);  // [na][cos, sin]
  int2* bounds = reinterpret_cast<int2*>(lds + layout.bounds);  // [2][kMaxAngles]: first row, last row
  W* win = reinterpret_cast<W*>(lds + layout.win);              // [2][max_rows][STRIDE]
  const int buffer = max_rows * STRIDE;

  const long long tile = xcd_tile(tiles);
  if (tile >= tiles) return;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int vl = wave * kVoxelsPerWave + lane / kLanesPerVoxel;  // tile slot of this lane's taps
  const int fl = lane % kLanesPerVoxel;                          // frame lane

  // Tile slot `lane`: this thread builds its taps for angles wave, wave + 8.
  float xr, zr, yr;
  const bool vr_ok = slot_voxel(p, map, tile, lane, xr, zr, yr) >= 0;
  // Steering cosines and sines once per workgroup (the float64 sincos is
  // register-hungry, so it runs before the accumulators are live); both
  // window buffers are zeroed once, so frames that are never loaded read 0.
  if (tid < p.na) {
    double sn, cs;
    sincos(static_cast<double>(p.angles[tid]), &sn, &cs);
    trig[2 * tid] = cs;
    trig[2 * tid + 1] = sn;
  }
  {
    uint4* w4 = reinterpret_cast<uint4*>(win);
    const int n = 2 * buffer * static_cast<int>(sizeof(W)) / 16;
    for (int i = tid; i < n; i += kThreads) w4[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  __syncthreads();

  // Channels run as one sequence q = aperture * nc + channel.  build(q): tap
  // table of the tile's voxels x angles for channel q (float64 receive and
  // transmit terms) and the exact row window of each angle.
  const int nq = 2 * p.nc;
  auto build = [&](int q) {
    const int cfg = q >= p.nc ? 1 : 0;
    const int ch = q - cfg * p.nc;
    Tap* tq = taps + (q & 1) * p.na * kTile;
    int2* bq = bounds + (q & 1) * kMaxAngles;
    Term r{kInactive, 0.f, 0.f, 0.f};
    if (vr_ok)
      r = receive_term(static_cast<double>(cfg == 0 ? xr : yr), static_cast<double>(zr),
                       static_cast<double>((cfg == 0 ? p.xe : p.ye)[ch]), p.c, p.fs, p.fd, p.fn);
    for (int a = wave; a < p.na; a += kWaves) {
      Tap tp{-1, 0.f, 0.f, 0.f};
      if (r.i != kInactive) {
        const Term t = transmit_term(static_cast<double>(cfg == 0 ? yr : xr), static_cast<double>(zr),
                                     trig[2 * a], trig[2 * a + 1], static_cast<double>(p.starts[a]), p.c, p.fs,
                                     p.fd);
        tp = combine(r, t, p.ns);
      }
      tq[a * kTile + lane] = tp;
      const bool live = tp.i0 >= 0;
      const int lo = wave_min(live ? tp.i0 : kBig);
      const int hi = wave_max(live ? tp.i0 + (tp.w != 0.f ? 1 : 0) : -kBig);
      if (lane == 0) bq[a] = make_int2(lo, hi);
    }
  };
  auto plane_of = [&](int q) {
    const int cfg = q >= p.nc ? 1 : 0;
    const W* data = static_cast<const W*>(cfg == 0 ? p.rc : p.cr);
    return data + static_cast<size_t>(q - cfg * p.nc) * p.na * p.ns * p.nt;
  };

  float2 acc[FPL];
#pragma unroll
  for (int j = 0; j < FPL; ++j) acc[j] = make_float2(0.f, 0.f);

  const size_t plane_size = static_cast<size_t>(p.ns) * p.nt;
  build(0);
  __syncthreads();
  // Software pipeline: while stage k is interpolated from one buffer, stage
  // k+1 (of this channel or, at a channel's end, the first of the next one)
  // streams into the other, and channel q+1's taps are built during channel
  // q's first stage.  One barrier per stage.
  int b = 0;
  Stage pending{0, 0, -1, 0, true, false};
  for (int q = 0; q < nq; ++q) {
    if (q + 1 < nq) build(q + 1);
    const Windows w = load_windows(bounds + (q & 1) * kMaxAngles, p.na);
    const W* plane0 = plane_of(q);
    Stage cur = pending;
    if (!cur.valid) {
      cur = plan_stage_lanes(w, 0, p.na, max_rows);
      if (cur.valid) issue_stage<W, STRIDE, kThreads>(plane0, plane_size, p.nt, cur, w, win + b * buffer, vec);
    }
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __syncthreads();  // stage landed; channel q+1's taps and windows visible
    pending.valid = false;
    const Tap* tq = taps + (q & 1) * p.na * kTile;
    while (cur.valid) {
      Stage nxt = next_stage_lanes(cur, w, p.na, max_rows);
      bool crossing = false;
      if (nxt.valid) {
        issue_stage<W, STRIDE, kThreads>(plane0, plane_size, p.nt, nxt, w, win + (b ^ 1) * buffer, vec);
      } else if (q + 1 < nq) {
        const Windows w1 = load_windows(bounds + ((q + 1) & 1) * kMaxAngles, p.na);
        nxt = plan_stage_lanes(w1, 0, p.na, max_rows);
        if (nxt.valid) issue_stage<W, STRIDE, kThreads>(plane_of(q + 1), plane_size, p.nt, nxt, w1, win + (b ^ 1) * buffer, vec);
        crossing = true;
      }
      const W* base = win + b * buffer + fl;
      int off = 0;
      for (int a = cur.a0; a < cur.a1; ++a) {
        int first;  // window row held in this buffer's row 0 for angle a
        bool take = true;
        const Tap tp = tq[a * kTile + vl];
        if (cur.r0 >= 0) {  // one chunk of a long window
          first = cur.r0;
          take = tp.i0 >= cur.r0 && (cur.last || tp.i0 < cur.r0 + max_rows - 1);
        } else {  // whole windows packed back to back
          const int2 bw = w.at(a);
          if (bw.x > bw.y) continue;
          first = bw.x - off;
          off += bw.y - bw.x + 1;
        }
        if (tp.i0 >= 0 && take) {
          const int i1 = tp.w != 0.f ? tp.i0 + 1 : tp.i0;
          accumulate<T, CP, FPL>(base + (tp.i0 - first) * STRIDE, base + (i1 - first) * STRIDE,
                                 make_coef<T, CP>(tp), acc);
        }
      }
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
      __syncthreads();
      b ^= 1;
      if (crossing) {
        pending = nxt;
        break;
      }
      cur = nxt;
    }
    if constexpr (APERTURES) {
      if (q == p.nc - 1 || q == nq - 1) {  // one aperture complete: store it, restart the sums
        const int cfg = q == nq - 1 ? 1 : 0;
        float xo, zo, yo;
        const long long v = slot_voxel(p, map, tile, vl, xo, zo, yo);
        if (v >= 0) {
#pragma unroll
          for (int j = 0; j < FPL; ++j) {
            const int t = fl + kLanesPerVoxel * j;
            if (t < p.nt) p.signal[(2 * v + cfg) * p.nt + t] = acc[j];
          }
          if (fl == 0) {
            p.pd[2 * v] = 0.f;
            p.pd[2 * v + 1] = 0.f;
          }
        }
#pragma unroll
        for (int j = 0; j < FPL; ++j) acc[j] = make_float2(0.f, 0.f);
      }
    }
  }
  if constexpr (!APERTURES) {
    float xo, zo, yo;
    const long long v = slot_voxel(p, map, tile, vl, xo, zo, yo);
    float power = 0.f;
#pragma unroll
    for (int j = 0; j < FPL; ++j) {
      const int t = fl + kLanesPerVoxel * j;
      if (t < p.nt && v >= 0) {
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
    if (fl == 0 && v >= 0) {
      p.pd[2 * v] = power / static_cast<float>(p.nt);
      p.pd[2 * v + 1] = 0.f;
    }
  }
}

template <typename T, int CP, int FPL, bool APERTURES>
hipError_t launch_shape(const EnsembleArgs& p, const TileMap& map, long long tiles, int out_storage) {
  using W = typename Word<T>::type;
  constexpr int FB = kLanesPerVoxel * FPL;
  constexpr int STRIDE = window_stride(FB, static_cast<int>(sizeof(W)));
  const int fixed = tile_layout(p.na, STRIDE, static_cast<int>(sizeof(W)), 0).bytes;
  const int max_rows = std::min(kMaxRows, (kLdsBytes - fixed) / (2 * STRIDE * static_cast<int>(sizeof(W))));
  if (max_rows < kMinRows) return hipErrorNotSupported;
  const TileLayout layout = tile_layout(p.na, STRIDE, static_cast<int>(sizeof(W)), max_rows);
  const long long grid = (tiles + 7) / 8 * 8;
  if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
  constexpr int U = 16 / static_cast<int>(sizeof(W));
  const bool vec = reinterpret_cast<uintptr_t>(p.rc) % 16 == 0 && reinterpret_cast<uintptr_t>(p.cr) % 16 == 0 &&
                   p.nt % U == 0;
  auto kernel = das_opw_tile<T, CP, FPL, APERTURES>;
  const hipError_t e = allow_dynamic_lds(reinterpret_cast<const void*>(kernel), kLdsBytes);
  if (e != hipSuccess) return e;
  kernel<<<dim3(static_cast<unsigned>(grid)), kThreads, layout.bytes>>>(p, map, tiles, max_rows, vec, out_storage);
  return hipGetLastError();
}

template <typename T, int CP, bool APERTURES>
hipError_t launch_frames(const EnsembleArgs& p, const TileMap& map, long long tiles, int out_storage) {
  if (p.nt <= 16) return launch_shape<T, CP, 2, APERTURES>(p, map, tiles, out_storage);
  if (p.nt <= 32) return launch_shape<T, CP, 4, APERTURES>(p, map, tiles, out_storage);
  if (p.nt <= 64) return launch_shape<T, CP, 8, APERTURES>(p, map, tiles, out_storage);
  if (p.nt <= 104) return launch_shape<T, CP, 13, APERTURES>(p, map, tiles, out_storage);
  return launch_shape<T, CP, 26, APERTURES>(p, map, tiles, out_storage);
}

hipError_t launch_any(const EnsembleArgs& p, const TileMap& map, long long tiles, int channel_storage,
                      int coefficients, int signal_storage, bool apertures) {
  if (p.na > kMaxAngles || p.nt > kLanesPerVoxel * kMaxFramesPerLane) return hipErrorNotSupported;
  // Plane offsets and in-plane row offsets are int.
  if (static_cast<long long>(p.ns) * p.nt > 0x7fffffffLL) return hipErrorNotSupported;
  if (signal_storage < RCABEAM_STORAGE_C64 || signal_storage > RCABEAM_STORAGE_BF16) return hipErrorInvalidValue;
  if (apertures) {
    if (channel_storage != RCABEAM_STORAGE_C64) return hipErrorInvalidValue;
    return launch_frames<float2, kCoefFloat, true>(p, map, tiles, signal_storage);
  }
  switch (channel_storage) {
    case RCABEAM_STORAGE_C64:
      return launch_frames<float2, kCoefFloat, false>(p, map, tiles, signal_storage);
    case RCABEAM_STORAGE_FP16:
      if (coefficients == kCoefSingle) return launch_frames<__half2, kCoefSingle, false>(p, map, tiles, signal_storage);
      if (coefficients == kCoefSplit) return launch_frames<__half2, kCoefSplit, false>(p, map, tiles, signal_storage);
      return launch_frames<__half2, kCoefFloat, false>(p, map, tiles, signal_storage);
    case RCABEAM_STORAGE_BF16:
      if (coefficients == kCoefSingle) return launch_frames<bf16x2, kCoefSingle, false>(p, map, tiles, signal_storage);
      return launch_frames<bf16x2, kCoefFloat, false>(p, map, tiles, signal_storage);
    default:
      return hipErrorInvalidValue;
  }
}

// ---- adjoint (transpose) of the OPW delay-and-sum ------------------------
//
// For IQ[v, t] = sum over taps of ((1 - w) S[i0, t] + w S[i0 + 1, t]) e^{i phi}
// and a real loss with gradient g = dL/dRe IQ + i dL/dIm IQ, the channel
// gradient is the adjoint applied to g: S[i0, t] += conj((1 - w) e^{i phi}) g,
// S[i0 + 1, t] += conj(w e^{i phi}) g, and
//   dL/dc = sum over taps, frames of Re(conj(g) dIQ/dc),
//   dIQ/dc = -(tau / c) e^{i phi} ((S[i0+1] - S[i0]) fs + i 2 pi fd v),
// v the interpolated sample, since every delay is a distance over c (the
// sample start t0 and the aperture mask do not depend on c).  The tiles and
// tap tables are the forward's (taps recomputed, nothing stored between the
// passes); frames go in chunks of at most 32.
//
// Voxels of a tile share sample rows, so their contributions collide.  Each
// half-wave (16 per workgroup) owns 32 frame columns and one private LDS
// gradient window: it takes its 4 voxels one after another, every lane adding
// to its own frame column with a plain read-modify-write (no LDS atomics, no
// address conflicts).  Per (channel, angle) chunk the 16 windows are summed
// and added to the packed gradient planes with global float atomics, or, in
// the deterministic form, as 64-bit fixed-point integers: integer addition is
// associative, so the order in which workgroups add does not matter.  Each
// window sum is rounded to a multiple of 2^-e, e chosen from max |grad IQ| so
// that no sample's total reaches 2^63, and a final pass converts the integers
// to float32 and adds them to the gradients.  dL/dc then goes to one partial
// per workgroup, summed in a fixed order.
struct AdjointArgs {
  EnsembleArgs p;     // rc, cr: packed complex64 channels of the forward pass
  const float2* giq;  // (nv, nt) gradient of the loss w.r.t. IQ
  float2* grc;        // packed (channels, angles, samples, frames) gradients, accumulated
  float2* gcr;
  double* gc;         // accumulated dL/dc
  int f0, nf;         // frame chunk [f0, f0 + nf) of this launch, nf <= 32
  unsigned long long* fixed;  // deterministic form: (re, im) int64 per gradient sample, RC planes then CR
  const unsigned* absmax;     // deterministic form: bits of max |component| of giq
  double* dc_partials;        // deterministic form: one dL/dc partial per workgroup of this launch
};

// Fixed-point exponent of the deterministic adjoint: every per-sample total is
// a sum of at most nv contributions of modulus <= |g| <= sqrt(2) m, so with
// m < 2^ex and nv <= 2^ln, 2^e nv sqrt(2) m < 2^62 for e = 61 - ln - ex.
__device__ __forceinline__ int fixed_exponent(const unsigned* absmax, long long nv) {
  const float m = __uint_as_float(absmax[0]);
  if (!(m > 0.f) || !(m < __builtin_inff())) return 0;
  int ex = 0;
  frexpf(m, &ex);
  int ln = 0;
  while ((1ll << ln) < nv) ++ln;
  return 61 - ln - ex;
}

constexpr int kHalf = kWave / 2;                      // frame columns of a half-wave
constexpr int kGroups = kThreads / kHalf;             // private gradient windows per workgroup
constexpr int kVoxelsPerGroup = kTile / kGroups;      // voxels a half-wave takes in turn

__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void das_opw_adjoint(
    AdjointArgs args, TileMap map, long long tiles, int max_rows) {
  const EnsembleArgs& p = args.p;
  extern __shared__ __align__(16) unsigned char lds[];
  Tap* taps = reinterpret_cast<Tap*>(lds);                                      // [na][kTile]
  double* trig = reinterpret_cast<double*>(lds + p.na * kTile * sizeof(Tap));   // [kMaxAngles][2]
  int2* bounds = reinterpret_cast<int2*>(trig + 2 * kMaxAngles);                // [kMaxAngles]
  double* partial = reinterpret_cast<double*>(bounds + kMaxAngles);             // [kWaves]
  float2* win_s = reinterpret_cast<float2*>(partial + kWaves);                  // [max_rows][kHalf] samples
  float2* win_g = win_s + max_rows * kHalf;                                     // [kGroups][max_rows][kHalf]

  const long long tile = xcd_tile(tiles);
  if (tile >= tiles) return;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int group = tid / kHalf;  // half-wave
  const int t = tid % kHalf;      // frame column (chunk-relative)
  float xr, zr, yr;
  const bool vr_ok = slot_voxel(p, map, tile, lane, xr, zr, yr) >= 0;
  if (tid < p.na) {
    double sn, cs;
    sincos(static_cast<double>(p.angles[tid]), &sn, &cs);
    trig[2 * tid] = cs;
    trig[2 * tid + 1] = sn;
  }
  // IQ gradients of this half-wave's voxels (slots group + 16 k) at frame t.
  float2 g[kVoxelsPerGroup];
#pragma unroll
  for (int k = 0; k < kVoxelsPerGroup; ++k) {
    float xo, zo, yo;
    const long long v = slot_voxel(p, map, tile, group + kGroups * k, xo, zo, yo);
    g[k] = (v >= 0 && t < args.nf) ? args.giq[v * p.nt + args.f0 + t] : make_float2(0.f, 0.f);
  }
  for (int i = tid; i < kGroups * max_rows * kHalf; i += kThreads) win_g[i] = make_float2(0.f, 0.f);
  float2* my_g = win_g + group * max_rows * kHalf + t;
  const int fe = args.fixed ? fixed_exponent(args.absmax, p.nv) : 0;
  const size_t aperture_size = static_cast<size_t>(p.nc) * p.na * p.ns * p.nt;
  const float fs = static_cast<float>(p.fs);
  const float omega = static_cast<float>(2.0 * 3.14159265358979323846 * p.fd);
  double dc = 0.0;
  const size_t plane_size = static_cast<size_t>(p.ns) * p.nt;

  for (int q = 0; q < 2 * p.nc; ++q) {
    const int cfg = q >= p.nc ? 1 : 0;
    const int ch = q - cfg * p.nc;
    __syncthreads();  // previous channel done with taps, bounds and windows
    {
      Term r{kInactive, 0.f, 0.f, 0.f};
      if (vr_ok)
        r = receive_term(static_cast<double>(cfg == 0 ? xr : yr), static_cast<double>(zr),
                         static_cast<double>((cfg == 0 ? p.xe : p.ye)[ch]), p.c, p.fs, p.fd, p.fn);
      for (int a = wave; a < p.na; a += kWaves) {
        Tap tp{-1, 0.f, 0.f, 0.f};
        if (r.i != kInactive) {
          const Term tt = transmit_term(static_cast<double>(cfg == 0 ? yr : xr), static_cast<double>(zr),
                                        trig[2 * a], trig[2 * a + 1], static_cast<double>(p.starts[a]), p.c, p.fs,
                                        p.fd);
          tp = combine(r, tt, p.ns);
        }
        taps[a * kTile + lane] = tp;
        const bool live = tp.i0 >= 0;
        const int lo = wave_min(live ? tp.i0 : kBig);
        const int hi = wave_max(live ? tp.i0 + (tp.w != 0.f ? 1 : 0) : -kBig);
        if (lane == 0) bounds[a] = make_int2(lo, hi);
      }
    }
    __syncthreads();
    const Windows wins = load_windows(bounds, p.na);
    const float2* plane0 = static_cast<const float2*>(cfg == 0 ? p.rc : p.cr) + static_cast<size_t>(ch) * p.na * plane_size;
    float2* gplane0 = (cfg == 0 ? args.grc : args.gcr) + static_cast<size_t>(ch) * p.na * plane_size;
    for (int a = 0; a < p.na; ++a) {
      const int2 bw = wins.at(a);
      if (bw.x > bw.y) continue;
      const float t0 = p.starts[a];
      for (int r0 = bw.x;; r0 += max_rows - 1) {
        const bool last = r0 + max_rows - 1 >= bw.y;
        const int rows = last ? bw.y + 1 - r0 : max_rows;
        __syncthreads();  // previous chunk's readers are done with win_s
        {
          const float2* src = plane0 + a * plane_size + static_cast<size_t>(r0) * p.nt + args.f0;
          for (int i = tid; i < rows * kHalf; i += kThreads) {
            const int row = i / kHalf;
            const int col = i - row * kHalf;
            win_s[i] = col < args.nf ? src[row * p.nt + col] : make_float2(0.f, 0.f);
          }
        }
        __syncthreads();
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < kVoxelsPerGroup; ++k) {
          const Tap tp = taps[a * kTile + group + kGroups * k];
          const bool in_chunk = tp.i0 >= 0 && tp.i0 >= r0 && (last || tp.i0 < r0 + max_rows - 1);
          if (in_chunk) {
            const int ra = tp.i0 - r0;
            const int rb = tp.w != 0.f ? ra + 1 : ra;
            const float u = 1.0f - tp.w;
            const float2 gk = g[k];
            const float2 c0 = cmulc(gk, make_float2(u * tp.c, u * tp.s));  // conj(P0) g
            const float2 c1 = cmulc(gk, make_float2(tp.w * tp.c, tp.w * tp.s));
            float2 wa = my_g[ra * kHalf];
            my_g[ra * kHalf] = make_float2(wa.x + c0.x, wa.y + c0.y);
            float2 wb = my_g[rb * kHalf];
            my_g[rb * kHalf] = make_float2(wb.x + c1.x, wb.y + c1.y);  // w == 0: rb == ra, c1 == 0
            const float2 s0 = win_s[ra * kHalf + t], s1 = win_s[rb * kHalf + t];
            const float2 d = make_float2(s1.x - s0.x, s1.y - s0.y);
            const float2 val = make_float2(fmaf(tp.w, d.x, s0.x), fmaf(tp.w, d.y, s0.y));
            // X = d fs + i omega val, Y = e^{i phi} X, Re(conj(g) Y), scaled by -tau / c
            const float2 x = make_float2(d.x * fs - omega * val.y, d.y * fs + omega * val.x);
            const float2 y = make_float2(tp.c * x.x - tp.s * x.y, tp.c * x.y + tp.s * x.x);
            const float tau = (static_cast<float>(tp.i0) + tp.w) / fs + t0;
            acc = fmaf(-tau, fmaf(gk.x, y.x, gk.y * y.y), acc);
          }
        }
        dc += static_cast<double>(acc) / p.c;
        __syncthreads();
        // Sum the 16 private windows, add to HBM, clear for the next chunk.
        float2* gdst = gplane0 + a * plane_size + static_cast<size_t>(r0) * p.nt + args.f0;
        const size_t fixed_base =
            cfg * aperture_size + (static_cast<size_t>(ch) * p.na + a) * plane_size + static_cast<size_t>(r0) * p.nt +
            args.f0;
        for (int i = tid; i < rows * kHalf; i += kThreads) {
          const int row = i / kHalf;
          const int col = i - row * kHalf;
          float2 sum = make_float2(0.f, 0.f);
#pragma unroll
          for (int gi = 0; gi < kGroups; ++gi) {
            float2* cell = win_g + (gi * max_rows + row) * kHalf + col;
            const float2 val = *cell;
            sum.x += val.x;
            sum.y += val.y;
            *cell = make_float2(0.f, 0.f);
          }
          if (col < args.nf) {
            if (args.fixed) {
              unsigned long long* cell = args.fixed + 2 * (fixed_base + static_cast<size_t>(row) * p.nt + col);
              if (sum.x != 0.f)
                atomicAdd(cell, static_cast<unsigned long long>(llrint(ldexp(static_cast<double>(sum.x), fe))));
              if (sum.y != 0.f)
                atomicAdd(cell + 1, static_cast<unsigned long long>(llrint(ldexp(static_cast<double>(sum.y), fe))));
            } else {
              if (sum.x != 0.f) unsafeAtomicAdd(&gdst[row * p.nt + col].x, sum.x);
              if (sum.y != 0.f) unsafeAtomicAdd(&gdst[row * p.nt + col].y, sum.y);
            }
          }
        }
        if (last) break;
      }
    }
  }
  // dL/dc: block sum in float64, one atomic per workgroup.
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) dc += __shfl_xor(dc, off, kWave);
  __syncthreads();
  if (lane == 0) partial[wave] = dc;
  __syncthreads();
  if (tid == 0) {
    double sum = 0.0;
    for (int i = 0; i < kWaves; ++i) sum += partial[i];
    if (args.dc_partials)
      args.dc_partials[blockIdx.x] = sum;
    else
      unsafeAtomicAdd(args.gc, sum);
  }
}

// Zero n 64-bit words (fixed-point planes and dL/dc partials) before a deterministic adjoint.
__global__ __launch_bounds__(256) void k_zero_words(unsigned long long* __restrict__ w, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) w[i] = 0ull;
}

// grad (RC planes then CR planes, float2) += fixed-point totals * 2^-e.
__global__ __launch_bounds__(256) void k_fixed_finalize(const unsigned long long* __restrict__ fixed,
                                                        const unsigned* __restrict__ absmax, long long nv,
                                                        long long per_aperture, float2* __restrict__ grc,
                                                        float2* __restrict__ gcr) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= 2 * per_aperture) return;
  const int fe = fixed_exponent(absmax, nv);
  const double re = ldexp(static_cast<double>(static_cast<long long>(fixed[2 * i])), -fe);
  const double im = ldexp(static_cast<double>(static_cast<long long>(fixed[2 * i + 1])), -fe);
  float2* g = i < per_aperture ? grc + i : gcr + (i - per_aperture);
  g->x += static_cast<float>(re);
  g->y += static_cast<float>(im);
}

// gc += sum of the dL/dc partials, in index order (one wave: fixed lane order, then a fixed-order butterfly).
__global__ __launch_bounds__(64) void k_dc_reduce(const double* __restrict__ partials, long long n, double* gc) {
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += 64) s += partials[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
  if (threadIdx.x == 0) *gc += s;
}

hipError_t launch_adjoint(const AdjointArgs& a, const TileMap& map, long long tiles) {
  const int fixed = a.p.na * kTile * static_cast<int>(sizeof(Tap)) + kMaxAngles * 16 + kMaxAngles * 8 + kWaves * 8;
  const int per_row = (1 + kGroups) * kHalf * 8;  // sample row + 16 gradient rows
  const int max_rows = std::min(kMaxRows, (kLdsBytes - fixed) / per_row);
  if (max_rows < kMinRows) return hipErrorNotSupported;
  const int bytes = fixed + max_rows * per_row;
  const long long grid = (tiles + 7) / 8 * 8;
  if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
  const hipError_t e = allow_dynamic_lds(reinterpret_cast<const void*>(das_opw_adjoint), kLdsBytes);
  if (e != hipSuccess) return e;
  das_opw_adjoint<<<dim3(static_cast<unsigned>(grid)), kThreads, bytes>>>(a, map, tiles, max_rows);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_opw_tile(const EnsembleArgs& p, int channel_storage, int coefficients, int signal_storage,
                           bool apertures) {
  const TileMap map{nullptr, nullptr, nullptr, 0, 0, 0, 0, false};
  return launch_any(p, map, (p.nv + kTile - 1) / kTile, channel_storage, coefficients, signal_storage, apertures);
}

hipError_t launch_opw_grid(const EnsembleArgs& p, const float* x, const float* z, const float* y, int nx, int nz,
                           int ny, int channel_storage, int coefficients, int signal_storage, bool apertures) {
  const int nbx = (nx + kPatch - 1) / kPatch, nby = (ny + kPatch - 1) / kPatch;
  const TileMap map{x, z, y, nx, nz, ny, nby, true};
  return launch_any(p, map, static_cast<long long>(nbx) * nby * nz, channel_storage, coefficients, signal_storage,
                    apertures);
}

}  // namespace rcb

namespace rcb {

size_t adjoint_workspace_bytes(const EnsembleArgs& p, int nx, int nz, int ny) {
  const int nbx = (nx + kPatch - 1) / kPatch, nby = (ny + kPatch - 1) / kPatch;
  const long long tiles = static_cast<long long>(nbx) * nby * nz;
  const long long grid = (tiles + 7) / 8 * 8;
  const long long chunks = (p.nt + kHalf - 1) / kHalf;
  const long long per_aperture = static_cast<long long>(p.nc) * p.na * p.ns * p.nt;
  return static_cast<size_t>(256 + (4 * per_aperture + chunks * grid) * 8);
}

hipError_t launch_opw_adjoint_grid(const EnsembleArgs& p, const float* x, const float* z, const float* y, int nx,
                                   int nz, int ny, const float2* giq, float2* grc, float2* gcr, double* gc,
                                   void* workspace) {
  if (p.na > kMaxAngles) return hipErrorNotSupported;
  if (static_cast<long long>(p.ns) * p.nt > 0x7fffffffLL) return hipErrorNotSupported;
  const int nbx = (nx + kPatch - 1) / kPatch, nby = (ny + kPatch - 1) / kPatch;
  const TileMap map{x, z, y, nx, nz, ny, nby, true};
  const long long tiles = static_cast<long long>(nbx) * nby * nz;
  const long long grid = (tiles + 7) / 8 * 8;
  const long long per_aperture = static_cast<long long>(p.nc) * p.na * p.ns * p.nt;
  // Deterministic form: workspace = [absmax (3 words, padded to 256 bytes)][fixed planes][dL/dc partials].
  unsigned* absmax = nullptr;
  unsigned long long* fixed = nullptr;
  double* partials = nullptr;
  const long long chunks = (p.nt + kHalf - 1) / kHalf;
  if (workspace) {
    absmax = static_cast<unsigned*>(workspace);
    fixed = reinterpret_cast<unsigned long long*>(static_cast<unsigned char*>(workspace) + 256);
    partials = reinterpret_cast<double*>(fixed + 4 * per_aperture);
    if (rcabeam_iq_absmax_device(giq, RCABEAM_STORAGE_C64, static_cast<size_t>(p.nv) * p.nt, absmax, 0) !=
        RCABEAM_SUCCESS)
      return hipErrorLaunchFailure;
    const long long words = 4 * per_aperture + chunks * grid;
    k_zero_words<<<static_cast<unsigned>((words + 255) / 256), 256>>>(fixed, words);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  // Frame chunks of at most 32: each chunk recomputes the taps (nothing is
  // stored between chunks or passes).
  for (int f0 = 0, k = 0; f0 < p.nt; f0 += kHalf, ++k) {
    const AdjointArgs a{p, giq, grc, gcr, gc, f0, std::min(kHalf, p.nt - f0), fixed, absmax,
                        partials ? partials + k * grid : nullptr};
    const hipError_t e = launch_adjoint(a, map, tiles);
    if (e != hipSuccess) return e;
  }
  if (workspace) {
    k_fixed_finalize<<<static_cast<unsigned>((2 * per_aperture + 255) / 256), 256>>>(fixed, absmax, p.nv,
                                                                                     per_aperture, grc, gcr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_dc_reduce<<<1, 64>>>(partials, chunks * grid, gc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace rcb
