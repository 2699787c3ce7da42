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
// kernel is HBM-bound in principle (0.5 ms per aperture of the full-size
// workload); its arithmetic, 2 K FMA per output, is close behind (~0.9 ms).
//
// Polyphase form (k_rf_to_iq_poly): with k = p D + q, output m is
//   sum_q sum_p g[p D + q] y_q[m + p],   y_q[n] = rf[n D + q - c],
// so a thread that computes R consecutive outputs of its column holds, per
// phase q, the R + P_q - 1 samples y_q[m0 .. m0 + R + P_q - 2] in registers
// (P_q = taps of phase q) and uses each of them for up to R outputs: about
// (R + K / D) D / R loads per output instead of K.  A workgroup is 256
// consecutive columns and R output rows; work items are dealt to XCDs in
// contiguous runs, output-row blocks fastest, so the blocks that share RF rows
// run on one XCD at about the same time and read them from its L2.  The output
// phasor depends on m only and is formed once per workgroup.  Filters with
// more than kPolyMaxTaps taps per phase use the direct kernel (k_rf_to_iq).
#include "common.hpp"
#include "rcabeam_mi355x.h"

namespace rcb {
namespace {

constexpr int kDemodThreads = 256;
constexpr int kDemodRows = 32;  // output rows per workgroup (direct kernel)
constexpr int kMaxTaps = 255;
constexpr int kPolyRows = 16;      // consecutive outputs per thread (polyphase kernel)
constexpr int kPolyMaxTaps = 32;   // most taps per phase the polyphase kernel takes

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

// Complex taps g[k] = 2 h[k] exp(-2j pi fd (k - c) / fs) into LDS (cycles reduced in float64).
__device__ __forceinline__ void form_taps(const DemodArgs& a, float2* g) {
  const int c = (a.ntaps - 1) / 2;
  for (int k = threadIdx.x; k < a.ntaps; k += kDemodThreads) {
    const double cycles = a.fd * static_cast<double>(k - c) / a.fs;
    const double r = cycles - rint(cycles);
    float sn, cs;
    sincospif(static_cast<float>(-2.0 * r), &sn, &cs);
    const float h2 = 2.0f * a.taps[k];
    g[k] = make_float2(h2 * cs, h2 * sn);
  }
}

// PMAX: taps per phase unrolled (>= every phase's count).  The window loads
// of a phase are issued together, unconditionally (rows outside the record or
// past the phase's window read a valid row and are zeroed afterwards; the
// test is wave-uniform), so a wave keeps R + PMAX - 1 loads in flight.
template <typename T, int PMAX>
__global__ __launch_bounds__(kDemodThreads) void k_rf_to_iq_poly(DemodArgs a, long long row_blocks,
                                                                 long long items) {
  __shared__ float2 g[kMaxTaps];
  __shared__ float2 phasor[kPolyRows];
  constexpr int W = kPolyRows + PMAX - 1;
  // Work item: (column block, output-row block), row blocks fastest, contiguous per XCD.
  const long long per_xcd = (items + 7) / 8;
  const long long item = static_cast<long long>(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (item >= items) return;  // whole workgroup: uniform
  const long long m0 = (item % row_blocks) * kPolyRows;
  const long long col = (item / row_blocks) * kDemodThreads + threadIdx.x;
  form_taps(a, g);
  if (threadIdx.x < kPolyRows) {
    // exp(-2j pi fd t_m), t_m = t0 + m D / fs
    const long long m = m0 + threadIdx.x;
    const double cycles = a.fd * (a.t0 + static_cast<double>(m * a.dec) / a.fs);
    const double r = cycles - rint(cycles);
    float sn, cs;
    sincospif(static_cast<float>(-2.0 * r), &sn, &cs);
    phasor[threadIdx.x] = make_float2(cs, sn);
  }
  __syncthreads();
  if (col >= a.cols) return;
  const T* base = static_cast<const T*>(a.rf) + col;
  const int c = (a.ntaps - 1) / 2;
  float2 acc[kPolyRows];
#pragma unroll
  for (int r = 0; r < kPolyRows; ++r) acc[r] = make_float2(0.f, 0.f);
  const long long first = m0 * a.dec - c;  // row of window element 0 of phase 0
  for (int q = 0; q < a.dec && q < a.ntaps; ++q) {
    const int taps_q = (a.ntaps - q + a.dec - 1) / a.dec;  // p D + q < K
    const int n_win = kPolyRows + taps_q - 1;
    T raw[W];
    uint64_t valid = 0;  // wave-uniform
    const long long n0 = first + q;
    if (n0 >= 0 && n0 + static_cast<long long>(W - 1) * a.dec < a.ns_rf) {
      // All W rows inside the record: a pointer stepped by D rows.
      const T* ptr = base + n0 * a.cols;
      const long long step = static_cast<long long>(a.dec) * a.cols;
#pragma unroll
      for (int i = 0; i < W; ++i) {
        raw[i] = *ptr;
        ptr += step;
      }
      valid = n_win >= 64 ? ~0ull : (1ull << n_win) - 1;
    } else {
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const long long n = n0 + static_cast<long long>(i) * a.dec;
        const bool ok = i < n_win && n >= 0 && n < a.ns_rf;
        valid |= ok ? (1ull << i) : 0ull;
        raw[i] = base[(ok ? n : 0ll) * a.cols];
      }
    }
    float win[W];
#pragma unroll
    for (int i = 0; i < W; ++i) win[i] = (valid >> i) & 1ull ? static_cast<float>(raw[i]) : 0.f;
#pragma unroll
    for (int p = 0; p < PMAX; ++p) {
      if (p < taps_q) {
        const float2 h = g[p * a.dec + q];
#pragma unroll
        for (int r = 0; r < kPolyRows; ++r) {
          acc[r].x = fmaf(h.x, win[r + p], acc[r].x);
          acc[r].y = fmaf(h.y, win[r + p], acc[r].y);
        }
      }
    }
  }
  const long long plane = col / a.nt, frame = col - plane * a.nt;
#pragma unroll
  for (int r = 0; r < kPolyRows; ++r) {
    const long long m = m0 + r;
    if (m >= a.ns_iq) break;
    const float2 e = phasor[r];
    const float2 v = make_float2(e.x * acc[r].x - e.y * acc[r].y, e.x * acc[r].y + e.y * acc[r].x);
    const long long index = a.packed ? (plane * a.ns_iq + m) * a.nt + frame : m * a.cols + col;
    a.out[index] = v;
  }
}

template <typename T>
void launch_poly(const DemodArgs& a, int taps_per_phase, long long row_blocks, long long items) {
  const unsigned grid = static_cast<unsigned>((items + 7) / 8 * 8);
  if (taps_per_phase <= 8)
    k_rf_to_iq_poly<T, 8><<<grid, kDemodThreads>>>(a, row_blocks, items);
  else if (taps_per_phase <= 16)
    k_rf_to_iq_poly<T, 16><<<grid, kDemodThreads>>>(a, row_blocks, items);
  else if (taps_per_phase <= 24)
    k_rf_to_iq_poly<T, 24><<<grid, kDemodThreads>>>(a, row_blocks, items);
  else
    k_rf_to_iq_poly<T, kPolyMaxTaps><<<grid, kDemodThreads>>>(a, row_blocks, items);
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
  const bool poly = (ntaps + decimation - 1) / decimation <= static_cast<size_t>(kPolyMaxTaps);
  const long long row_blocks = (static_cast<long long>(ns_iq) + kPolyRows - 1) / kPolyRows;
  const long long items = gx * row_blocks;
  if (poly ? (items + 7) / 8 * 8 > 0x7fffffffLL : (gx > 0x7fffffffLL || gy > 65535)) return RCABEAM_ERROR_ARGUMENT;
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
  if (poly) {
    const int per_phase = static_cast<int>((ntaps + decimation - 1) / decimation);
    if (rf_type == 0)
      launch_poly<float>(a, per_phase, row_blocks, items);
    else
      launch_poly<short>(a, per_phase, row_blocks, items);
    return launch_status();
  }
  const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>(gy));
  if (rf_type == 0)
    k_rf_to_iq<float><<<grid, kDemodThreads>>>(a);
  else
    k_rf_to_iq<short><<<grid, kDemodThreads>>>(a);
  return launch_status();
}
