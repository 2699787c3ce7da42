// RF -> IQ demodulation on MI355X: mixing with the carrier, low-pass FIR and
// decimation of real RF channel samples into the complex IQ the beamformer
// takes.  The reference starts from IQ and lists RF-to-IQ processing as open
// work (reference ROADMAP.md:58); its IQ convention is fixed by its simulator
// (src/rcabeam/sim.py:188: an echo delayed by tau carries exp(-2j pi fd tau))
// and its delay-and-sum (exp(+2j pi fd tau), sim.py:385).  For real RF
// rf(t) = env(t - tau) cos(2 pi fd (t - tau)) this is
//   iq(t) = LPF{ 2 rf(t) exp(-2j pi fd t) } = env(t - tau) exp(-2j pi fd tau).
//
// Output sample m sits at t_m = t0 + m D / fs_rf (the RF start time, IQ rate
// fs_rf / D) and, with the centred FIR h of K (odd) taps, c = (K - 1) / 2,
//   iq[m] = sum_k 2 h[k] rf[m D + k - c] exp(-2j pi fd t_{mD+k-c})
//         = exp(-2j pi fd t_m) sum_k g[k] rf[m D + k - c],
//   g[k]  = 2 h[k] exp(-2j pi fd (k - c) / fs_rf),
// samples outside [0, ns_rf) being zero.  The complex taps g are formed once
// per workgroup in float64 and kept in LDS; the output phasor exp(-2j pi fd
// t_m) is reduced in float64 and evaluated in float32 (as common.hpp does for
// delays).  Layout: RF and host-layout IQ are (samples, channels, angles,
// frames); a thread owns one (channel, angle, frame) column, so every load and
// store of a wave is 64 consecutive columns of one sample row.  The packed
// output is the frame-inner (channels, angles, samples, frames) layout the
// delay-and-sum kernels read, so demodulation also does the packing.  The
// kernel is HBM-bound: each RF row is fetched from HBM once and re-read K / D
// times from cache by the output rows of a block.
#include "common.hpp"
#include "rcabeam_mi355x.h"

namespace rcb {
namespace {

constexpr int kDemodThreads = 256;
constexpr int kDemodRows = 32;  // output rows per workgroup
constexpr int kMaxTaps = 255;

struct DemodArgs {
  const void* rf;
  const float* taps;
  float2* out;
  long long ns_rf, ns_iq, cols;  // cols = nc * na * nt
  int ntaps, dec, nt, packed;
  double t0, fs, fd;
};

template <typename T>
__device__ __forceinline__ float rf_value(const T* rf, long long index);
template <>
__device__ __forceinline__ float rf_value<float>(const float* rf, long long index) {
  return rf[index];
}
template <>
__device__ __forceinline__ float rf_value<short>(const short* rf, long long index) {
  return static_cast<float>(rf[index]);
}

template <typename T>
__global__ __launch_bounds__(kDemodThreads) void k_rf_to_iq(DemodArgs a) {
  __shared__ float2 g[kMaxTaps];
  const int c = (a.ntaps - 1) / 2;
  for (int k = threadIdx.x; k < a.ntaps; k += kDemodThreads) {
    // exp(-2j pi fd (k - c) / fs): cycles reduced to [-1/2, 1/2] in float64.
    const double cycles = a.fd * static_cast<double>(k - c) / a.fs;
    const double r = cycles - rint(cycles);
    float sn, cs;
    sincospif(static_cast<float>(-2.0 * r), &sn, &cs);
    const float h2 = 2.0f * a.taps[k];
    g[k] = make_float2(h2 * cs, h2 * sn);
  }
  __syncthreads();
  const long long col = static_cast<long long>(blockIdx.x) * kDemodThreads + threadIdx.x;
  if (col >= a.cols) return;
  const T* rf = static_cast<const T*>(a.rf);
  const long long m0 = static_cast<long long>(blockIdx.y) * kDemodRows;
  const long long m1 = m0 + kDemodRows < a.ns_iq ? m0 + kDemodRows : a.ns_iq;
  const long long plane = col / a.nt, frame = col - plane * a.nt;
  for (long long m = m0; m < m1; ++m) {
    const long long n0 = m * a.dec - c;
    float re = 0.f, im = 0.f;
    if (n0 >= 0 && n0 + a.ntaps <= a.ns_rf) {
      const T* src = rf + n0 * a.cols + col;
      for (int k = 0; k < a.ntaps; ++k) {
        const float v = rf_value(src, static_cast<long long>(k) * a.cols);
        re = fmaf(g[k].x, v, re);
        im = fmaf(g[k].y, v, im);
      }
    } else {
      for (int k = 0; k < a.ntaps; ++k) {
        const long long n = n0 + k;
        if (n < 0 || n >= a.ns_rf) continue;
        const float v = rf_value(rf, n * a.cols + col);
        re = fmaf(g[k].x, v, re);
        im = fmaf(g[k].y, v, im);
      }
    }
    // exp(-2j pi fd t_m), t_m = t0 + m D / fs
    const double cycles = a.fd * (a.t0 + static_cast<double>(m * a.dec) / a.fs);
    const double r = cycles - rint(cycles);
    float sn, cs;
    sincospif(static_cast<float>(-2.0 * r), &sn, &cs);
    const float2 v = make_float2(cs * re - sn * im, cs * im + sn * re);
    const long long index = a.packed ? (plane * a.ns_iq + m) * a.nt + frame : m * a.cols + col;
    a.out[index] = v;
  }
}

}  // namespace
}  // namespace rcb

using namespace rcb;

extern "C" rcabeam_status rcabeam_rf_to_iq_device(const void* rf, int rf_type, size_t ns_rf, size_t nc, size_t na,
                                                  size_t nt, const float* taps, size_t ntaps, size_t decimation,
                                                  float fs_rf, float f_demod, float t_start, void* iq, size_t ns_iq,
                                                  int packed) {
  if (!rf || !taps || !iq || !ns_rf || !nc || !na || !nt || !ns_iq || !decimation || ntaps % 2 == 0 ||
      ntaps > static_cast<size_t>(kMaxTaps) || (rf_type != 0 && rf_type != 1))
    return RCABEAM_ERROR_ARGUMENT;
  if (!(fs_rf > 0.f) || !(fs_rf < __builtin_inff()) || !(f_demod == f_demod) || !(t_start == t_start) ||
      f_demod - f_demod != 0.f || t_start - t_start != 0.f)
    return RCABEAM_ERROR_ARGUMENT;
  if (decimation > 0x7fffffff || nt > 0x7fffffff || ns_iq > (ns_rf + decimation - 1) / decimation)
    return RCABEAM_ERROR_ARGUMENT;
  const long long cols = static_cast<long long>(nc) * static_cast<long long>(na) * static_cast<long long>(nt);
  const long long gx = (cols + kDemodThreads - 1) / kDemodThreads;
  const long long gy = (static_cast<long long>(ns_iq) + kDemodRows - 1) / kDemodRows;
  if (gx > 0x7fffffffLL || gy > 65535) return RCABEAM_ERROR_ARGUMENT;
  DemodArgs a;
  a.rf = rf;
  a.taps = taps;
  a.out = static_cast<float2*>(iq);
  a.ns_rf = static_cast<long long>(ns_rf);
  a.ns_iq = static_cast<long long>(ns_iq);
  a.cols = cols;
  a.ntaps = static_cast<int>(ntaps);
  a.dec = static_cast<int>(decimation);
  a.nt = static_cast<int>(nt);
  a.packed = packed ? 1 : 0;
  a.t0 = t_start;
  a.fs = fs_rf;
  a.fd = f_demod;
  const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>(gy));
  if (rf_type == 0)
    k_rf_to_iq<float><<<grid, kDemodThreads>>>(a);
  else
    k_rf_to_iq<short><<<grid, kDemodThreads>>>(a);
  return launch_status();
}
