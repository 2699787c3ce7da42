// Synthetic RCA channel acquisitions on the GPU.
//
// Point scatterers with the Gaussian-envelope echo model of reference
// src/rcabeam/sim.py:131-189: sample time t_s = t0 + s / fs, delay
// tau = (z cos a + u sin a) / c + sqrt(z^2 + (v - e)^2) / c with (v, u) = (x, y)
// for RC and (y, x) for CR, echo amp * exp(-((t_s - tau) / sigma)^2) *
// exp(-2j pi fd tau).  The envelope is evaluated within +/- 6 sigma of tau
// (exp(-36) ~ 2e-16 beyond).  Delays and phases are float64; the echo sum is
// float32 in LDS.
//
// Scene model for frame t: static scatterers carry a per-group complex
// temporal factor g[t][group]; moving scatterers carry their own position per
// frame; circular complex Gaussian noise (counter-based hash, seeded) is added
// per sample.  One workgroup owns one (channel, angle) pair of one aperture:
// it builds the static rows once in LDS, then produces frames in chunks and
// writes them in the (samples, channels, angles, frames) layout, frames
// innermost, so each store run covers a chunk of consecutive frames.
#include "common.hpp"
#include "rcabeam_mi355x.h"

namespace rcb {

constexpr int kSceneThreads = 1024;
constexpr int kSceneChunk = 16;

struct SceneArgs {
  const double* static_pos;     // (n_static, 3): x, z, y
  const double* static_amp;     // (n_static, 2)
  const int* static_group;      // (n_static)
  const double* group_factor;   // (nt, n_groups, 2)
  const double* moving_pos;     // (nt, n_moving, 3)
  const double* moving_amp;     // (n_moving, 2)
  const float* elements;        // (nc)
  const float* angles;          // (na)
  float2* out;                  // (ns, nc, na, nt)
  long long n_static, n_moving;
  int n_groups, ns, nc, na, nt, config;
  double t0, c, fs, fd, sigma, noise_std;
  unsigned long long seed;
};

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Two independent standard normals from one 64-bit counter (Box-Muller).
__device__ __forceinline__ float2 gaussian_pair(unsigned long long seed, unsigned long long counter) {
  const unsigned long long h = mix64(seed ^ mix64(counter));
  const double u1 = (static_cast<double>(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = static_cast<double>(mix64(h) >> 11) * (1.0 / 9007199254740992.0);
  const double r = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  return make_float2(static_cast<float>(r * c), static_cast<float>(r * s));
}

// Add one echo to an LDS row of ns complex samples (interleaved float pairs).
__device__ __forceinline__ void add_echo(float* row, const SceneArgs& p, double x, double z, double y, double e,
                                         double cos_a, double sin_a, double ar, double ai) {
  const double v = p.config == 0 ? x : y;
  const double u = p.config == 0 ? y : x;
  const double tau = (z * cos_a + u * sin_a) / p.c + sqrt(z * z + (v - e) * (v - e)) / p.c;
  double sn, cs;
  sincospi(-2.0 * (p.fd * tau - floor(p.fd * tau)), &sn, &cs);
  const double er = ar * cs - ai * sn, ei = ar * sn + ai * cs;
  const double center = (tau - p.t0) * p.fs;
  const double half = 6.0 * p.sigma * p.fs;
  int s0 = static_cast<int>(ceil(center - half));
  int s1 = static_cast<int>(floor(center + half));
  if (s0 < 0) s0 = 0;
  if (s1 > p.ns - 1) s1 = p.ns - 1;
  for (int s = s0; s <= s1; ++s) {
    const double d = (p.t0 + s / p.fs - tau) / p.sigma;
    const float env = static_cast<float>(exp(-d * d));
    atomicAdd(row + 2 * s, env * static_cast<float>(er));
    atomicAdd(row + 2 * s + 1, env * static_cast<float>(ei));
  }
}

__global__ __launch_bounds__(kSceneThreads) void k_scene(SceneArgs p) {
  extern __shared__ __align__(16) float lds[];
  float* stat = lds;                                      // [n_groups][ns][2]
  float* frames = lds + 2 * p.n_groups * p.ns;            // [kSceneChunk][ns][2]
  const int ch = blockIdx.x % p.nc;
  const int a = blockIdx.x / p.nc;
  const double e = p.elements[ch];
  double sin_a, cos_a;
  sincos(static_cast<double>(p.angles[a]), &sin_a, &cos_a);
  for (int i = threadIdx.x; i < 2 * p.n_groups * p.ns; i += blockDim.x) stat[i] = 0.f;
  __syncthreads();
  for (long long k = threadIdx.x; k < p.n_static; k += blockDim.x) {
    const double* q = p.static_pos + 3 * k;
    add_echo(stat + 2 * static_cast<long long>(p.static_group[k]) * p.ns, p, q[0], q[1], q[2], e, cos_a, sin_a,
             p.static_amp[2 * k], p.static_amp[2 * k + 1]);
  }
  __syncthreads();
  for (int f0 = 0; f0 < p.nt; f0 += kSceneChunk) {
    const int nf = p.nt - f0 < kSceneChunk ? p.nt - f0 : kSceneChunk;
    for (int i = threadIdx.x; i < 2 * kSceneChunk * p.ns; i += blockDim.x) frames[i] = 0.f;
    __syncthreads();
    for (long long k = threadIdx.x; k < p.n_moving * nf; k += blockDim.x) {
      const int f = static_cast<int>(k / p.n_moving);
      const long long m = k % p.n_moving;
      const double* q = p.moving_pos + (static_cast<long long>(f0 + f) * p.n_moving + m) * 3;
      add_echo(frames + 2 * f * p.ns, p, q[0], q[1], q[2], e, cos_a, sin_a, p.moving_amp[2 * m],
               p.moving_amp[2 * m + 1]);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < p.ns * kSceneChunk; i += blockDim.x) {
      const int s = i / kSceneChunk, f = i % kSceneChunk;
      if (f >= nf) continue;
      const int t = f0 + f;
      float vr = frames[2 * (f * p.ns + s)], vi = frames[2 * (f * p.ns + s) + 1];
      for (int g = 0; g < p.n_groups; ++g) {
        const double* gf = p.group_factor + (static_cast<long long>(t) * p.n_groups + g) * 2;
        const float sr = stat[2 * (g * p.ns + s)], si = stat[2 * (g * p.ns + s) + 1];
        vr += static_cast<float>(gf[0]) * sr - static_cast<float>(gf[1]) * si;
        vi += static_cast<float>(gf[0]) * si + static_cast<float>(gf[1]) * sr;
      }
      const long long index = ((static_cast<long long>(s) * p.nc + ch) * p.na + a) * p.nt + t;
      if (p.noise_std > 0) {
        const float2 n = gaussian_pair(p.seed + 0x51ull * static_cast<unsigned long long>(p.config + 1),
                                       static_cast<unsigned long long>(index));
        vr += static_cast<float>(p.noise_std) * n.x;
        vi += static_cast<float>(p.noise_std) * n.y;
      }
      p.out[index] = make_float2(vr, vi);
    }
    __syncthreads();
  }
}

}  // namespace rcb

using namespace rcb;

extern "C" size_t rcabeam_simulate_lds_bytes(size_t ns, int n_groups) {
  return sizeof(float) * 2 * ns * (static_cast<size_t>(n_groups) + kSceneChunk);
}

extern "C" rcabeam_status rcabeam_simulate_channels_device(
    const double* static_pos, const double* static_amp, const int* static_group, size_t n_static,
    const double* group_factor, int n_groups, const double* moving_pos, const double* moving_amp, size_t n_moving,
    const float* elements, const float* angles, void* out, size_t ns, size_t nc, size_t na, size_t nt, int config,
    float t_start, float c, float fs, float fd, float sigma_t, float noise_std, unsigned long long seed) {
  if (!out || !elements || !angles || !ns || !nc || !na || !nt || (config != 0 && config != 1)) return RCABEAM_ERROR_ARGUMENT;
  if (n_static && (!static_pos || !static_amp || !static_group || !group_factor || n_groups < 1)) return RCABEAM_ERROR_ARGUMENT;
  if (n_moving && (!moving_pos || !moving_amp)) return RCABEAM_ERROR_ARGUMENT;
  if (!(c > 0) || !(fs > 0) || !(sigma_t > 0) || !(noise_std >= 0)) return RCABEAM_ERROR_ARGUMENT;
  if (nc * na > 0x7fffffffULL || ns > 0x7fffffff || nt > 0x7fffffff) return RCABEAM_ERROR_ARGUMENT;
  const int groups = n_static ? n_groups : 0;
  const size_t lds = rcabeam_simulate_lds_bytes(ns, groups);
  if (lds > 160 * 1024) return RCABEAM_ERROR_ARGUMENT;
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_scene), hipFuncAttributeMaxDynamicSharedMemorySize,
                          static_cast<int>(lds)) != hipSuccess)
    return RCABEAM_ERROR_CUDA;
  SceneArgs p;
  p.static_pos = static_pos;
  p.static_amp = static_amp;
  p.static_group = static_group;
  p.group_factor = group_factor;
  p.moving_pos = moving_pos;
  p.moving_amp = moving_amp;
  p.elements = elements;
  p.angles = angles;
  p.out = static_cast<float2*>(out);
  p.n_static = static_cast<long long>(n_static);
  p.n_moving = static_cast<long long>(n_moving);
  p.n_groups = groups;
  p.ns = static_cast<int>(ns);
  p.nc = static_cast<int>(nc);
  p.na = static_cast<int>(na);
  p.nt = static_cast<int>(nt);
  p.config = config;
  p.t0 = t_start;
  p.c = c;
  p.fs = fs;
  p.fd = fd;
  p.sigma = sigma_t;
  p.noise_std = noise_std;
  p.seed = seed;
  k_scene<<<dim3(static_cast<unsigned>(nc * na)), kSceneThreads, lds>>>(p);
  return launch_status();
}
