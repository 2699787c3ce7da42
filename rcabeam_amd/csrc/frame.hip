// Single-frame entry points of the C ABI (reference src/rcabeam/core.cu):
// per-angle / per-channel DAS of one aperture, the fused one-frame powers, and
// the combiners on pre-beamformed per-angle volumes.  Channels use the host
// layout (samples, channels, angles); taps use the same float64-derived
// receive/transmit terms as the ensemble kernels (common.hpp).  These paths
// are latency-sized (one frame), so one thread owns one output element.
#include "common.hpp"
#include "rcabeam_mi355x.h"

namespace rcb {

struct FrameGeom {
  const float* scan;
  const float* angles;
  const float* starts;
  int ns, nc, na;
  long long nv;
  double c, fs, fd, fn;
};

__device__ __forceinline__ Term frame_rx(const FrameGeom& g, const float* el, long long v, int ch, int config) {
  const double x = g.scan[3 * v], z = g.scan[3 * v + 1], y = g.scan[3 * v + 2];
  return receive_term(config == 0 ? x : y, z, static_cast<double>(el[ch]), g.c, g.fs, g.fd, g.fn);
}

__device__ __forceinline__ Term frame_tx(const FrameGeom& g, long long v, int a, int config) {
  const double x = g.scan[3 * v], z = g.scan[3 * v + 1], y = g.scan[3 * v + 2];
  double sn, cs;
  sincos(static_cast<double>(g.angles[a]), &sn, &cs);
  return transmit_term(config == 0 ? y : x, z, cs, sn, static_cast<double>(g.starts[a]), g.c, g.fs, g.fd);
}

__device__ __forceinline__ float2 frame_sample(const float2* iq, const FrameGeom& g, int ch, int a, const Term& r,
                                               const Term& t) {
  const Tap tp = combine(r, t, g.ns);
  if (tp.i0 < 0) return make_float2(0.f, 0.f);
  const int i1 = tp.i0 == g.ns - 1 ? tp.i0 : tp.i0 + 1;
  const float2 A = iq[(static_cast<size_t>(tp.i0) * g.nc + ch) * g.na + a];
  const float2 B = iq[(static_cast<size_t>(i1) * g.nc + ch) * g.na + a];
  const float vx = fmaf(tp.w, B.x - A.x, A.x), vy = fmaf(tp.w, B.y - A.y, A.y);
  return make_float2(tp.c * vx - tp.s * vy, tp.c * vy + tp.s * vx);
}

// Sum over channels of one aperture and one angle.
__device__ float2 angle_sum(const float2* iq, const float* el, const FrameGeom& g, long long v, int a, int config) {
  const Term t = frame_tx(g, v, a, config);
  float2 acc = make_float2(0.f, 0.f);
  for (int ch = 0; ch < g.nc; ++ch) acc = cadd(acc, frame_sample(iq, g, ch, a, frame_rx(g, el, v, ch, config), t));
  return acc;
}

__global__ void k_angle_sums(const float2* iq, const float* el, FrameGeom g, int config, float2* out) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= g.nv * g.na) return;
  const long long v = i / g.na;
  const int a = static_cast<int>(i - v * g.na);
  out[i] = angle_sum(iq, el, g, v, a, config);
}

__global__ void k_channel_sums(const float2* iq, const float* el, FrameGeom g, int config, float2* out) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= g.nv * g.nc) return;
  const long long v = i / g.nc;
  const int ch = static_cast<int>(i - v * g.nc);
  const Term r = frame_rx(g, el, v, ch, config);
  float2 acc = make_float2(0.f, 0.f);
  for (int a = 0; a < g.na; ++a) acc = cadd(acc, frame_sample(iq, g, ch, a, r, frame_tx(g, v, a, config)));
  out[i] = acc;
}

__device__ float active_count(const float* el, const FrameGeom& g, long long v, int config) {
  float n = 0.f;
  for (int ch = 0; ch < g.nc; ++ch) n += frame_rx(g, el, v, ch, config).i == kInactive ? 0.f : 1.f;
  return n;
}

__global__ void k_active(const float* el, FrameGeom g, int config, float* out) {
  const long long v = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < g.nv) out[v] = active_count(el, g, v, config);
}

enum FusedMode { kFusedOpw = 0, kFusedXdoppler = 1, kFusedFmas = 2, kFusedAll = 3 };

__global__ void k_fused(const float2* rc, const float2* cr, const float* xe, const float* ye, FrameGeom g, int mode,
                        float* out0, float* out1, float* out2) {
  const long long v = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= g.nv) return;
  float2 r = make_float2(0.f, 0.f), q = r, fr = r, fq = r;
  for (int a = 0; a < g.na; ++a) {
    const float2 ar = angle_sum(rc, xe, g, v, a, 0);
    const float2 aq = angle_sum(cr, ye, g, v, a, 1);
    r = cadd(r, ar);
    q = cadd(q, aq);
    fr = cadd(fr, sgnsqrt(ar));
    fq = cadd(fq, sgnsqrt(aq));
  }
  const float opw = cnorm(cadd(r, q));
  const float xd = cabs(cmulc(r, q));
  const float fm = cnorm(cmul(fr, fq));
  if (mode == kFusedOpw) out0[v] = opw;
  if (mode == kFusedXdoppler) out0[v] = xd;
  if (mode == kFusedFmas) out0[v] = fm;
  if (mode == kFusedAll) {
    out0[v] = opw;
    out1[v] = xd;
    out2[v] = fm;
  }
}

// One-frame DMAS-CCF-ACF (reference src/rcabeam/core.cu:537-627).
__global__ void k_dmas(const float2* rc, const float2* cr, const float* xe, const float* ye, FrameGeom g,
                       float* out) {
  const long long v = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= g.nv) return;
  float2 sqrt_sum = make_float2(0.f, 0.f);
  float abs_sum = 0.f;
  for (int ch = 0; ch < g.nc; ++ch) {
    const Term rr = frame_rx(g, xe, v, ch, 0), rq = frame_rx(g, ye, v, ch, 1);
    float2 sr = make_float2(0.f, 0.f), sq = sr;
    for (int a = 0; a < g.na; ++a) {
      sr = cadd(sr, frame_sample(rc, g, ch, a, rr, frame_tx(g, v, a, 0)));
      sq = cadd(sq, frame_sample(cr, g, ch, a, rq, frame_tx(g, v, a, 1)));
    }
    sqrt_sum = cadd(sqrt_sum, cadd(sgnsqrt(sr), sgnsqrt(sq)));
    abs_sum += cabs(sr) + cabs(sq);
  }
  const float y = cnorm(sqrt_sum) - abs_sum;
  const float norm = active_count(xe, g, v, 0) + active_count(ye, g, v, 1);
  const float wc = (abs_sum > 0.f && norm > 0.f) ? fmaxf(y, 0.f) / norm / abs_sum : 0.f;
  float2 sr = make_float2(0.f, 0.f), sq = sr;
  float er = 0.f, eq = 0.f;
  for (int a = 0; a < g.na; ++a) {
    const float2 ar = angle_sum(rc, xe, g, v, a, 0);
    const float2 aq = angle_sum(cr, ye, g, v, a, 1);
    sr = cadd(sr, ar);
    sq = cadd(sq, aq);
    er += cnorm(ar);
    eq += cnorm(aq);
  }
  const float inv = 1.0f / static_cast<float>(g.na);
  const float2 mr = make_float2(sr.x * inv, sr.y * inv), mq = make_float2(sq.x * inv, sq.y * inv);
  const float nr = cnorm(mr), nq = cnorm(mq);
  const float var = fmaxf(er * inv - nr, 0.f) + fmaxf(eq * inv - nq, 0.f);
  const float num = cabs(cmulc(mr, mq));
  const float den = nr + nq + 0.5f * var - num;
  const float wa = den > 0.f ? num / den : 0.f;
  const float result = y * wc * wa;
  out[v] = isfinite(result) ? result : 0.f;
}

// ---- combiners on pre-beamformed per-angle volumes (reference core.cu:48-107)
__global__ void k_opw_angles(const float2* iq, float2* out, long long nv, long long na, long long nt) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nv * nt) return;
  const long long v = i / nt, t = i - v * nt;
  float2 acc = make_float2(0.f, 0.f);
  for (long long a = 0; a < na; ++a) acc = cadd(acc, iq[(v * na + a) * nt + t]);
  out[i] = acc;
}

__global__ void k_span_pd(const float2* iq, float* out, long long nv, long long na, long long nt, long long rc0,
                          long long rcn, long long cr0, long long crn, int fmas) {
  const long long v = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const float2* base = iq + v * na * nt;
  float2 mean = make_float2(0.f, 0.f);
  float power = 0.f;
  for (long long t = 0; t < nt; ++t) {
    float2 r = make_float2(0.f, 0.f), q = r;
    for (long long a = 0; a < rcn; ++a) {
      const float2 x = base[(rc0 + a) * nt + t];
      r = cadd(r, fmas == 1 ? sgnsqrt(x) : x);
    }
    for (long long a = 0; a < crn; ++a) {
      const float2 x = base[(cr0 + a) * nt + t];
      q = cadd(q, fmas == 1 ? sgnsqrt(x) : x);
    }
    if (fmas == 1)
      power += cnorm(cmul(r, q));
    else
      mean = cadd(mean, cmulc(r, q));
  }
  const float inv = 1.0f / static_cast<float>(nt);
  if (fmas == 2) {  // complex mean of the XDoppler signal
    out[2 * v] = mean.x * inv;
    out[2 * v + 1] = mean.y * inv;
    return;
  }
  out[v] = fmas ? power * inv : hypotf(mean.x * inv, mean.y * inv);
}

static bool bad_frame_args(const void* iq, const float* scan, const float* el, const float* angles,
                           const float* starts, const void* out, size_t ns, size_t nc, size_t na, size_t nv,
                           int config) {
  return !iq || !scan || !el || !angles || !starts || !out || ns < 2 || !nc || !na || !nv ||
         (config != 0 && config != 1) || ns > 0x7fffffff || nc > 0x7fffffff || na > 0x7fffffff;
}

static FrameGeom make_geom(const float* scan, const float* angles, const float* starts, size_t ns, size_t nc,
                           size_t na, size_t nv, float c, float fs, float fd, float fn) {
  return FrameGeom{scan, angles, starts, static_cast<int>(ns), static_cast<int>(nc), static_cast<int>(na),
                   static_cast<long long>(nv), c, fs, fd, fn};
}

static unsigned blocks_for(long long n, int threads) { return static_cast<unsigned>((n + threads - 1) / threads); }

static rcabeam_status fused(const void* rc, const void* cr, const float* scan, const float* xe, const float* ye,
                            const float* angles, const float* starts, float* o0, float* o1, float* o2, size_t ns,
                            size_t nc, size_t na, size_t nv, float c, float fs, float fd, float fn, int mode) {
  if (bad_frame_args(rc, scan, xe, angles, starts, o0, ns, nc, na, nv, 0) || !cr || !ye) return RCABEAM_ERROR_ARGUMENT;
  if (mode == kFusedAll && (!o1 || !o2)) return RCABEAM_ERROR_ARGUMENT;
  const FrameGeom g = make_geom(scan, angles, starts, ns, nc, na, nv, c, fs, fd, fn);
  k_fused<<<blocks_for(g.nv, 128), 128>>>(static_cast<const float2*>(rc), static_cast<const float2*>(cr), xe, ye, g,
                                          mode, o0, o1, o2);
  return launch_status();
}

}  // namespace rcb

using namespace rcb;

extern "C" const char* rcabeam_status_string(rcabeam_status status) {
  switch (status) {
    case RCABEAM_SUCCESS:
      return "success";
    case RCABEAM_ERROR_CUDA:
      return "device runtime error";
    case RCABEAM_ERROR_ARGUMENT:
      return "invalid argument";
  }
  return "unknown error";
}

extern "C" rcabeam_status rcabeam_opw_device(const void* iq, void* out, size_t n_voxels, size_t n_angles,
                                             size_t n_frames) {
  if (!iq || !out || !n_voxels || !n_angles || !n_frames) return RCABEAM_ERROR_ARGUMENT;
  const long long n = static_cast<long long>(n_voxels) * static_cast<long long>(n_frames);
  k_opw_angles<<<blocks_for(n, 256), 256>>>(static_cast<const float2*>(iq), static_cast<float2*>(out),
                                            static_cast<long long>(n_voxels), static_cast<long long>(n_angles),
                                            static_cast<long long>(n_frames));
  return launch_status();
}

static rcabeam_status span_pd(const void* iq, float* out, size_t nv, size_t na, size_t nt, size_t rc0, size_t rcn,
                              size_t cr0, size_t crn, int fmas) {
  if (!iq || !out || !nv || !na || !nt) return RCABEAM_ERROR_ARGUMENT;
  if (rc0 + rcn > na || cr0 + crn > na) return RCABEAM_ERROR_ARGUMENT;
  k_span_pd<<<blocks_for(static_cast<long long>(nv), 256), 256>>>(
      static_cast<const float2*>(iq), out, static_cast<long long>(nv), static_cast<long long>(na),
      static_cast<long long>(nt), static_cast<long long>(rc0), static_cast<long long>(rcn),
      static_cast<long long>(cr0), static_cast<long long>(crn), fmas);
  return launch_status();
}

extern "C" rcabeam_status rcabeam_xdoppler_pd_device(const void* iq, float* out, size_t n_voxels, size_t n_angles,
                                                     size_t n_frames, size_t rc_start, size_t rc_count,
                                                     size_t cr_start, size_t cr_count) {
  return span_pd(iq, out, n_voxels, n_angles, n_frames, rc_start, rc_count, cr_start, cr_count, 0);
}

extern "C" rcabeam_status rcabeam_xdoppler_mean_device(const void* iq, float* out, size_t n_voxels, size_t n_angles,
                                                       size_t n_frames, size_t rc_start, size_t rc_count,
                                                       size_t cr_start, size_t cr_count) {
  return span_pd(iq, out, n_voxels, n_angles, n_frames, rc_start, rc_count, cr_start, cr_count, 2);
}

extern "C" rcabeam_status rcabeam_rc_fmas_pd_device(const void* iq, float* out, size_t n_voxels, size_t n_angles,
                                                    size_t n_frames, size_t rc_start, size_t rc_count,
                                                    size_t cr_start, size_t cr_count) {
  return span_pd(iq, out, n_voxels, n_angles, n_frames, rc_start, rc_count, cr_start, cr_count, 1);
}

extern "C" rcabeam_status rcabeam_delay_rca_channels_device(const void* iq_ch, const float* scan_coords_m,
                                                            const float* elements_m, const float* angles_rad,
                                                            const float* t_start_s, void* out, size_t n_samples,
                                                            size_t n_channels, size_t n_angles, size_t n_voxels,
                                                            int config, float sound_speed_m_s,
                                                            float sampling_freq_hz, float demod_freq_hz,
                                                            float f_number) {
  if (bad_frame_args(iq_ch, scan_coords_m, elements_m, angles_rad, t_start_s, out, n_samples, n_channels, n_angles,
                     n_voxels, config))
    return RCABEAM_ERROR_ARGUMENT;
  const FrameGeom g = make_geom(scan_coords_m, angles_rad, t_start_s, n_samples, n_channels, n_angles, n_voxels,
                                sound_speed_m_s, sampling_freq_hz, demod_freq_hz, f_number);
  k_angle_sums<<<blocks_for(g.nv * g.na, 128), 128>>>(static_cast<const float2*>(iq_ch), elements_m, g, config,
                                                      static_cast<float2*>(out));
  return launch_status();
}

extern "C" rcabeam_status rcabeam_delay_rca_channel_data_device(
    const void* iq_ch, const float* scan_coords_m, const float* elements_m, const float* angles_rad,
    const float* t_start_s, void* channel_sum, void* angle_sum, float* n_active, size_t n_samples,
    size_t n_channels, size_t n_angles, size_t n_voxels, int config, float sound_speed_m_s, float sampling_freq_hz,
    float demod_freq_hz, float f_number) {
  if (bad_frame_args(iq_ch, scan_coords_m, elements_m, angles_rad, t_start_s, angle_sum, n_samples, n_channels,
                     n_angles, n_voxels, config) ||
      !channel_sum || !n_active)
    return RCABEAM_ERROR_ARGUMENT;
  const FrameGeom g = make_geom(scan_coords_m, angles_rad, t_start_s, n_samples, n_channels, n_angles, n_voxels,
                                sound_speed_m_s, sampling_freq_hz, demod_freq_hz, f_number);
  const float2* iq = static_cast<const float2*>(iq_ch);
  k_angle_sums<<<blocks_for(g.nv * g.na, 128), 128>>>(iq, elements_m, g, config, static_cast<float2*>(angle_sum));
  k_channel_sums<<<blocks_for(g.nv * g.nc, 128), 128>>>(iq, elements_m, g, config,
                                                        static_cast<float2*>(channel_sum));
  k_active<<<blocks_for(g.nv, 128), 128>>>(elements_m, g, config, n_active);
  return launch_status();
}

extern "C" rcabeam_status rcabeam_opw_pd_from_channels_device(
    const void* iq_rc, const void* iq_cr, const float* scan_coords_m, const float* x_elements_m,
    const float* y_elements_m, const float* angles_rad, const float* t_start_s, float* out, size_t n_samples,
    size_t n_channels, size_t n_angles, size_t n_voxels, float sound_speed_m_s, float sampling_freq_hz,
    float demod_freq_hz, float f_number) {
  return fused(iq_rc, iq_cr, scan_coords_m, x_elements_m, y_elements_m, angles_rad, t_start_s, out, nullptr,
               nullptr, n_samples, n_channels, n_angles, n_voxels, sound_speed_m_s, sampling_freq_hz, demod_freq_hz,
               f_number, kFusedOpw);
}

extern "C" rcabeam_status rcabeam_xdoppler_pd_from_channels_device(
    const void* iq_rc, const void* iq_cr, const float* scan_coords_m, const float* x_elements_m,
    const float* y_elements_m, const float* angles_rad, const float* t_start_s, float* out, size_t n_samples,
    size_t n_channels, size_t n_angles, size_t n_voxels, float sound_speed_m_s, float sampling_freq_hz,
    float demod_freq_hz, float f_number) {
  return fused(iq_rc, iq_cr, scan_coords_m, x_elements_m, y_elements_m, angles_rad, t_start_s, out, nullptr,
               nullptr, n_samples, n_channels, n_angles, n_voxels, sound_speed_m_s, sampling_freq_hz, demod_freq_hz,
               f_number, kFusedXdoppler);
}

extern "C" rcabeam_status rcabeam_rc_fmas_pd_from_channels_device(
    const void* iq_rc, const void* iq_cr, const float* scan_coords_m, const float* x_elements_m,
    const float* y_elements_m, const float* angles_rad, const float* t_start_s, float* out, size_t n_samples,
    size_t n_channels, size_t n_angles, size_t n_voxels, float sound_speed_m_s, float sampling_freq_hz,
    float demod_freq_hz, float f_number) {
  return fused(iq_rc, iq_cr, scan_coords_m, x_elements_m, y_elements_m, angles_rad, t_start_s, out, nullptr,
               nullptr, n_samples, n_channels, n_angles, n_voxels, sound_speed_m_s, sampling_freq_hz, demod_freq_hz,
               f_number, kFusedFmas);
}

extern "C" rcabeam_status rcabeam_fast_pd_from_channels_device(
    const void* iq_rc, const void* iq_cr, const float* scan_coords_m, const float* x_elements_m,
    const float* y_elements_m, const float* angles_rad, const float* t_start_s, float* opw_out, float* xdoppler_out,
    float* rc_fmas_out, size_t n_samples, size_t n_channels, size_t n_angles, size_t n_voxels,
    float sound_speed_m_s, float sampling_freq_hz, float demod_freq_hz, float f_number) {
  return fused(iq_rc, iq_cr, scan_coords_m, x_elements_m, y_elements_m, angles_rad, t_start_s, opw_out,
               xdoppler_out, rc_fmas_out, n_samples, n_channels, n_angles, n_voxels, sound_speed_m_s,
               sampling_freq_hz, demod_freq_hz, f_number, kFusedAll);
}

extern "C" rcabeam_status rcabeam_dmas_ccf_acf_from_channels_device(
    const void* iq_rc, const void* iq_cr, const float* scan_coords_m, const float* x_elements_m,
    const float* y_elements_m, const float* angles_rad, const float* t_start_s, float* out, size_t n_samples,
    size_t n_channels, size_t n_angles, size_t n_voxels, float sound_speed_m_s, float sampling_freq_hz,
    float demod_freq_hz, float f_number) {
  if (bad_frame_args(iq_rc, scan_coords_m, x_elements_m, angles_rad, t_start_s, out, n_samples, n_channels,
                     n_angles, n_voxels, 0) ||
      !iq_cr || !y_elements_m)
    return RCABEAM_ERROR_ARGUMENT;
  const FrameGeom g = make_geom(scan_coords_m, angles_rad, t_start_s, n_samples, n_channels, n_angles, n_voxels,
                                sound_speed_m_s, sampling_freq_hz, demod_freq_hz, f_number);
  k_dmas<<<blocks_for(g.nv, 128), 128>>>(static_cast<const float2*>(iq_rc), static_cast<const float2*>(iq_cr),
                                         x_elements_m, y_elements_m, g, out);
  return launch_status();
}
