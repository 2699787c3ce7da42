// OPW delay-and-sum on MFMA for 16-bit (FP16 / BF16) channel storage, on a
// Cartesian (x, z, y) grid.  Mixed-precision form of the reference OPW kernel
// (src/rcabeam/ensemble.cu:341-454): float32 accumulation, 16-bit samples.
//
// For one aperture, angle and channel, the contribution of a voxel is
//   sum_t-independent:  out[v, t] += c0_v X[i0_v, t] + c1_v X[i0_v + 1, t],
// with c0 = (1 - w) e^{i phi}, c1 = w e^{i phi}.  For 16 voxels that form a
// 4 x 4 lateral patch at one depth, all sample rows i0_v, i0_v + 1 of one
// (aperture, channel, angle) fall in a window of a few rows (<= 8 for the
// usual geometries; wider spans take more windows), so the update is the
// small dense product
//   OUT (16 voxels x 32 frames, complex) += C (16 x 8 rows) * X (8 rows x 32 frames)
// written in real form on v_mfma_f32_32x32x16_{f16,bf16}:
//   M = 32 = (voxel, re/im of the output), K = 16 = (row, re/im of the sample),
//   N = 32 frames.
// The sample window is read straight from the packed (channels, angles,
// samples, frames) layout: each lane loads 4 rows x 1 frame of 4-byte
// complex samples, 32 consecutive frames per row across the wave.  The
// coefficients are split into hi + lo 16-bit parts (two MFMAs), so their
// rounding (~2^-22) is negligible next to the 16-bit storage of the samples,
// whose products with the coefficients are exact in the float32 accumulator.
//
// One wave owns one 16-voxel tile and all frames (in chunks of 32 * FB);
// receive terms are float64 per (voxel, channel), transmit terms float64 per
// (voxel, angle) in LDS, as in the VALU kernels (common.hpp).
#include "common.hpp"
#include "rcabeam_mi355x.h"

namespace rcb {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kMfmaBlock = 256;  // 4 waves = 4 tiles
constexpr int kTileWaves = kMfmaBlock / kWave;

struct GridArgs {
  const void* rc;
  const void* cr;
  const float* x;
  const float* z;
  const float* y;
  const float* xe;
  const float* ye;
  const float* angles;
  const float* starts;
  float* pd;
  void* signal;
  int nx, nz, ny;
  int ns, nc, na, nt;
  int tx4, ty4;  // tiles along x and y
  double c, fs, fd, fn;
};

// 16-bit storage traits: how a 4-byte sample becomes two 16-bit MFMA operands.
template <typename T>
struct Half;
template <>
struct Half<__half2> {
  typedef f16x8 vec;
  typedef _Float16 elem;
  static __device__ __forceinline__ unsigned short bits(float f) { return __half_as_ushort(__float2half_rn(f)); }
  static __device__ __forceinline__ float value(unsigned short b) { return __half2float(__ushort_as_half(b)); }
  static __device__ __forceinline__ floatx16 mfma(vec a, vec b, floatx16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Half<bf16x2> {
  typedef bf16x8_t vec;
  typedef __bf16 elem;
  static __device__ __forceinline__ unsigned short bits(float f) { return static_cast<unsigned short>(bf16_bits_rne(f)); }
  static __device__ __forceinline__ float value(unsigned short b) { return __uint_as_float(static_cast<uint32_t>(b) << 16); }
  static __device__ __forceinline__ floatx16 mfma(vec a, vec b, floatx16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};

template <typename T, typename O, int FB>
__global__ __launch_bounds__(kMfmaBlock) void das_opw_mfma(GridArgs p, long long tiles) {
  typedef Half<T> H;
  typedef typename H::vec vec;
  typedef typename H::elem elem;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  Term* txl = reinterpret_cast<Term*>(lds_raw) + static_cast<long long>(wave) * 16 * p.na;

  // XCD-aware block order, then one tile per wave.
  const long long blocks = (tiles + kTileWaves - 1) / kTileWaves;
  const long long per_xcd = (blocks + 7) / 8;
  const long long block = (static_cast<long long>(blockIdx.x) % 8) * per_xcd + blockIdx.x / 8;
  const long long tile = block * kTileWaves + wave;
  const bool tile_ok = block < blocks && tile < tiles;

  // Tile -> (ix4, iz, iy4); lane -> voxel of the A operand (row of the output).
  long long rest = tile_ok ? tile : 0;
  const int iy4 = static_cast<int>(rest % p.ty4);
  rest /= p.ty4;
  const int iz = static_cast<int>(rest % p.nz);
  const int ix4 = static_cast<int>(rest / p.nz);
  const int m = lane & 31;
  const int vm = m >> 1;  // voxel 0..15 of the tile (x-major 4 x 4)
  const int q = m & 1;    // 0: real row of the output, 1: imaginary row
  const int kg = lane >> 5;
  const int ixm = 4 * ix4 + (vm >> 2), iym = 4 * iy4 + (vm & 3);
  const bool vm_ok = tile_ok && ixm < p.nx && iym < p.ny;
  const double xm = p.x[ixm < p.nx ? ixm : p.nx - 1];
  const double ym = p.y[iym < p.ny ? iym : p.ny - 1];
  const double zm = p.z[iz];

  float pw[4][2] = {};
  for (int f0 = 0; f0 < p.nt; f0 += 32 * FB) {
    floatx16 acc[FB];
#pragma unroll
    for (int b = 0; b < FB; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[b][e] = 0.f;
    int fcl[FB];
#pragma unroll
    for (int b = 0; b < FB; ++b) {
      const int f = f0 + 32 * b + m;
      fcl[b] = f < p.nt ? f : p.nt - 1;
    }
    for (int cfg = 0; cfg < 2; ++cfg) {
      const T* __restrict__ data = static_cast<const T*>(cfg == 0 ? p.rc : p.cr);
      const float* el = cfg == 0 ? p.xe : p.ye;
      __syncthreads();
      for (int idx = lane; idx < 16 * p.na; idx += kWave) {
        const int v = idx / p.na, a = idx - v * p.na;
        const int ix = 4 * ix4 + (v >> 2), iy = 4 * iy4 + (v & 3);
        const double xv = p.x[ix < p.nx ? ix : p.nx - 1], yv = p.y[iy < p.ny ? iy : p.ny - 1];
        double sn, cs;
        sincos(static_cast<double>(p.angles[a]), &sn, &cs);
        txl[idx] = transmit_term(cfg == 0 ? yv : xv, zm, cs, sn, static_cast<double>(p.starts[a]), p.c, p.fs, p.fd);
      }
      __syncthreads();
      for (int ch = 0; ch < p.nc; ++ch) {
        const Term rx = vm_ok ? receive_term(cfg == 0 ? xm : ym, zm, static_cast<double>(el[ch]), p.c, p.fs, p.fd, p.fn)
                              : Term{kInactive, 0.f, 0.f, 0.f};
        for (int a = 0; a < p.na; ++a) {
          const Tap tp = combine(rx, txl[vm * p.na + a], p.ns);
          const bool on = tp.i0 >= 0;
          int lo_row = on ? tp.i0 : 0x7fffffff;
          int hi_row = on ? tp.i0 + (tp.w > 0.f ? 1 : 0) : -1;
#pragma unroll
          for (int off = 2; off < 32; off <<= 1) {
            lo_row = min(lo_row, __shfl_xor(lo_row, off, kWave));
            hi_row = max(hi_row, __shfl_xor(hi_row, off, kWave));
          }
          lo_row = uniform(lo_row);
          hi_row = uniform(hi_row);
          if (hi_row < lo_row) continue;  // no voxel of the tile sees this (channel, angle)
          const float c0r = (1.f - tp.w) * tp.c, c0i = (1.f - tp.w) * tp.s;
          const float c1r = tp.w * tp.c, c1i = tp.w * tp.s;
          const size_t plane = (static_cast<size_t>(ch) * p.na + a) * p.ns;
          for (int wb = lo_row; wb <= hi_row; wb += 8) {
            // A operand: rows wb + 4*kg + j (j = 0..3), (re, im) of the sample for each.
            vec ahi, alo;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int r = wb + 4 * kg + j;
              float cr_ = 0.f, ci_ = 0.f;
              if (on && r == tp.i0) {
                cr_ = c0r;
                ci_ = c0i;
              } else if (on && r == tp.i0 + 1) {
                cr_ = c1r;
                ci_ = c1i;
              }
              // out_re += Cr Xr - Ci Xi ; out_im += Ci Xr + Cr Xi
              const float a_re = q == 0 ? cr_ : ci_;
              const float a_im = q == 0 ? -ci_ : cr_;
              const unsigned short hr = H::bits(a_re), hi_ = H::bits(a_im);
              const unsigned short lr = H::bits(a_re - H::value(hr)), li = H::bits(a_im - H::value(hi_));
              ahi[2 * j] = __builtin_bit_cast(elem, hr);
              ahi[2 * j + 1] = __builtin_bit_cast(elem, hi_);
              alo[2 * j] = __builtin_bit_cast(elem, lr);
              alo[2 * j + 1] = __builtin_bit_cast(elem, li);
            }
            const T* rowp[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              int r = wb + 4 * kg + j;
              r = r < p.ns ? r : p.ns - 1;
              rowp[j] = data + (plane + r) * p.nt;
            }
#pragma unroll
            for (int b = 0; b < FB; ++b) {
              uint32_t w4[4];
#pragma unroll
              for (int j = 0; j < 4; ++j) w4[j] = *reinterpret_cast<const uint32_t*>(rowp[j] + fcl[b]);
              vec bv;
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                bv[2 * j] = __builtin_bit_cast(elem, static_cast<unsigned short>(w4[j] & 0xffffu));
                bv[2 * j + 1] = __builtin_bit_cast(elem, static_cast<unsigned short>(w4[j] >> 16));
              }
              acc[b] = H::mfma(ahi, bv, acc[b]);
              acc[b] = H::mfma(alo, bv, acc[b]);
            }
          }
        }
      }
    }
    // Epilogue: lane holds frame f0 + 32 b + m for output rows i = 8 g + 4 kg + (0..3),
    // i.e. (re, im) of voxels 4 g + 2 kg + h, h = 0, 1.
#pragma unroll
    for (int b = 0; b < FB; ++b) {
      const int f = f0 + 32 * b + m;
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int v = 4 * g + 2 * kg + h;
          const int ix = 4 * ix4 + (v >> 2), iy = 4 * iy4 + (v & 3);
          if (!tile_ok || ix >= p.nx || iy >= p.ny || f >= p.nt) continue;
          const float2 val = make_float2(acc[b][4 * g + 2 * h], acc[b][4 * g + 2 * h + 1]);
          pw[g][h] += cnorm(val);
          if (p.signal) {
            const long long id = (static_cast<long long>(ix) * p.nz + iz) * p.ny + iy;
            store_sample(static_cast<O*>(p.signal) + id * p.nt + f, val);
          }
        }
    }
  }
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float s = pw[g][h];
#pragma unroll
      for (int off = 1; off < 32; off <<= 1) s += __shfl_xor(s, off, kWave);
      const int v = 4 * g + 2 * kg + h;
      const int ix = 4 * ix4 + (v >> 2), iy = 4 * iy4 + (v & 3);
      if (m == 0 && tile_ok && ix < p.nx && iy < p.ny) {
        const long long id = (static_cast<long long>(ix) * p.nz + iz) * p.ny + iy;
        p.pd[2 * id] = s / static_cast<float>(p.nt);
        p.pd[2 * id + 1] = 0.f;
      }
    }
}

template <typename T, typename O>
static hipError_t launch_mfma(const GridArgs& p) {
  const long long tiles = static_cast<long long>(p.tx4) * p.nz * p.ty4;
  const long long blocks = (tiles + kTileWaves - 1) / kTileWaves;
  const long long grid = ((blocks + 7) / 8) * 8;
  const size_t lds = static_cast<size_t>(kTileWaves) * 16 * p.na * sizeof(Term);
  if (grid > 0x7fffffffLL || lds > 64 * 1024) return hipErrorInvalidValue;
  das_opw_mfma<T, O, 7><<<dim3(static_cast<unsigned>(grid)), kMfmaBlock, lds>>>(p, tiles);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_mfma_out(const GridArgs& p, int signal_storage) {
  switch (signal_storage) {
    case RCABEAM_STORAGE_C64:
      return launch_mfma<T, float2>(p);
    case RCABEAM_STORAGE_FP16:
      return launch_mfma<T, __half2>(p);
    case RCABEAM_STORAGE_BF16:
      return launch_mfma<T, bf16x2>(p);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace rcb

using namespace rcb;

extern "C" rcabeam_status rcabeam_opw_grid_device(const void* rc, const void* cr, int channel_storage, const float* x,
                                                  const float* z, const float* y, size_t nx, size_t nz, size_t ny,
                                                  const float* xe, const float* ye, const float* angles,
                                                  const float* starts, float* pd, void* signal, int signal_storage,
                                                  size_t ns, size_t nc, size_t na, size_t nt, float c, float fs,
                                                  float fd, float fn) {
  if (!rc || !cr || !x || !z || !y || !xe || !ye || !angles || !starts || !pd) return RCABEAM_ERROR_ARGUMENT;
  if (!nx || !nz || !ny || ns < 2 || !nc || !na || !nt) return RCABEAM_ERROR_ARGUMENT;
  if (nx > 0x7fffffff || nz > 0x7fffffff || ny > 0x7fffffff || ns > 0x7fffffff || nc > 0x7fffffff ||
      na > 0x7fffffff || nt > 0x7fffffff)
    return RCABEAM_ERROR_ARGUMENT;
  if (!(c > 0) || !(fs > 0) || c != c || fs != fs || fd != fd || fn != fn || c - c != 0 || fs - fs != 0 ||
      fd - fd != 0 || fn - fn != 0)
    return RCABEAM_ERROR_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(rc) % 4 || reinterpret_cast<uintptr_t>(cr) % 4) return RCABEAM_ERROR_ARGUMENT;
  GridArgs p;
  p.rc = rc;
  p.cr = cr;
  p.x = x;
  p.z = z;
  p.y = y;
  p.xe = xe;
  p.ye = ye;
  p.angles = angles;
  p.starts = starts;
  p.pd = pd;
  p.signal = signal;
  p.nx = static_cast<int>(nx);
  p.nz = static_cast<int>(nz);
  p.ny = static_cast<int>(ny);
  p.ns = static_cast<int>(ns);
  p.nc = static_cast<int>(nc);
  p.na = static_cast<int>(na);
  p.nt = static_cast<int>(nt);
  p.tx4 = static_cast<int>((nx + 3) / 4);
  p.ty4 = static_cast<int>((ny + 3) / 4);
  p.c = c;
  p.fs = fs;
  p.fd = fd;
  p.fn = fn;
  hipError_t e;
  switch (channel_storage) {
    case RCABEAM_STORAGE_FP16:
      e = launch_mfma_out<__half2>(p, signal_storage);
      break;
    case RCABEAM_STORAGE_BF16:
      e = launch_mfma_out<bf16x2>(p, signal_storage);
      break;
    default:
      return RCABEAM_ERROR_ARGUMENT;
  }
  if (e == hipErrorInvalidValue) return RCABEAM_ERROR_ARGUMENT;
  return e == hipSuccess ? RCABEAM_SUCCESS : RCABEAM_ERROR_CUDA;
}
