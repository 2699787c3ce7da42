// Hybrid DMAS-CCF-ACF approximation (reference examples/_support/dmas_delay.cu:21-151,
// driven by examples/_support/dmas_delay.py:61-154): delay tables for the two
// marginals DMAS-CCF-ACF needs, instead of per-channel and per-angle
// delay-and-sum.
//
// * Channel tables (per depth slab): C[(u, ch, k), t] = sum over angles of
//   the sample at the slab's steering anchor + k * step, rotated by the
//   anchor's carrier phase.  One wave per (u, ch, k), lanes over frames.
// * Angle planes: A[(v, k, a), t] = sum over aperture channels of the sample
//   at the receive term r(v, ch, k) + transmit term (k, a), rotated by the
//   baseband excess phase; a not-a-knot spline along k follows (the two-stage
//   spline entry point).  One wave per (v, k, a), lanes over frames.
// * Finish: one wave per voxel, lanes over frames (two per lane, one
//   16-byte load, for an even frame count).  Each active channel's
//   excess receive delay indexes its C row (linear, or Catmull-Rom inside
//   the table), the rotated values give the signed-root DMAS sums and the CCF
//   magnitudes; each angle's spline lookup of A, rotated by the residual
//   carrier, gives the ACF sums; DMAS x CCF x ACF is averaged over frames.
//   Voxels run depth-fastest within a slab so consecutive waves read the same
//   table rows at nearby knots, and each XCD takes a contiguous run of them.
//
// Tables are float32 (x, y, z, w) quadruples as the caller forms them; all
// reads of packed channels (channels, angles, samples, frames) are rows of
// contiguous frames, so every wave load is one coalesced 512-byte access.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.hpp"
#include "rcabeam_mi355x.h"

namespace rcb {
namespace {

constexpr int kHybridWaves = 4;  // waves per 256-thread workgroup
constexpr int kXcdCount = 8;

__device__ __forceinline__ float2 rotate(float2 x, float c, float s) {
  return make_float2(x.x * c - x.y * s, x.x * s + x.y * c);
}
__device__ __forceinline__ float2 lerp2(float2 a, float2 b, float w) {
  return make_float2(fmaf(w, b.x - a.x, a.x), fmaf(w, b.y - a.y, a.y));
}
__device__ __forceinline__ float2 add2(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 scale2(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// F consecutive frames from p (F = 2: one 16-byte load; the caller keeps p
// 16-byte aligned).
template <int F>
__device__ __forceinline__ void load_frames(const float2* p, float2 (&v)[F]) {
  if constexpr (F == 1) {
    v[0] = p[0];
  } else {
    const float4 q = *reinterpret_cast<const float4*>(p);
    v[0] = make_float2(q.x, q.y);
    v[1] = make_float2(q.z, q.w);
  }
}

// Frames [t, t + F) of a (ns, nt) trace, linearly interpolated at a
// wave-uniform position and rotated by (c, s), added to acc; nothing outside
// the closed window [0, ns - 1].
template <int F>
__device__ __forceinline__ void add_sample(const float2* trace, float pos, int ns, int nt, int t, float c, float s,
                                           float2 (&acc)[F]) {
  if (!(pos >= 0.f) || pos > static_cast<float>(ns - 1)) return;
  const int i = static_cast<int>(pos);
  float2 a[F], b[F];
  load_frames<F>(trace + static_cast<size_t>(i) * nt + t, a);
  load_frames<F>(trace + static_cast<size_t>(min(i + 1, ns - 1)) * nt + t, b);
#pragma unroll
  for (int f = 0; f < F; ++f) acc[f] = add2(acc[f], rotate(lerp2(a[f], b[f], pos - static_cast<float>(i)), c, s));
}

// Store frames [t, t + F) that exist.
template <int F>
__device__ __forceinline__ void store_frames(float2* out, int t, int nt, const float2 (&v)[F]) {
  if (t >= nt) return;
  if constexpr (F == 1)
    out[t] = v[0];
  else
    *reinterpret_cast<float4*>(out + t) = make_float4(v[0].x, v[0].y, v[1].x, v[1].y);
}

// Block b's logical index: each XCD (blocks b % 8) takes a contiguous run.
__device__ __forceinline__ long long xcd_block(long long blocks) {
  const long long per = (blocks + kXcdCount - 1) / kXcdCount;
  return static_cast<long long>(blockIdx.x % kXcdCount) * per + blockIdx.x / kXcdCount;
}

template <int F>
__global__ __launch_bounds__(256) void k_hybrid_planes(const float2* __restrict__ raw, const float4* __restrict__ prx,
                                                       const float* __restrict__ ptx, float2* __restrict__ out,
                                                       long long entries, int nc, int na, int ns, int nt, int nk) {
  const long long entry = static_cast<long long>(blockIdx.x) * kHybridWaves + (threadIdx.x >> 6);
  if (entry >= entries) return;
  const int lane = threadIdx.x & 63;
  const int a = static_cast<int>(entry % na);
  const int k = static_cast<int>((entry / na) % nk);
  const long long v = entry / na / nk;
  const float tx = ptx[k * na + a];
  for (int t0 = 0; t0 < nt; t0 += 64 * F) {
    const int t = t0 + F * lane;
    const int tt = t < nt ? t : nt - F;
    float2 sum[F];
#pragma unroll
    for (int f = 0; f < F; ++f) sum[f] = make_float2(0.f, 0.f);
    for (int ch = 0; ch < nc; ++ch) {
      const float4 r = prx[(v * nc + ch) * nk + k];
      if (r.w == 0.f) continue;
      const float2* trace = raw + (static_cast<size_t>(ch) * na + a) * ns * static_cast<size_t>(nt);
      add_sample<F>(trace, r.x + tx, ns, nt, tt, r.y, r.z, sum);
    }
    store_frames<F>(out + entry * nt, t, nt, sum);
  }
}

template <int F>
__global__ __launch_bounds__(256) void k_hybrid_tables(const float2* __restrict__ raw,
                                                       const float4* __restrict__ anchors, float2* __restrict__ out,
                                                       long long entries, int ns, int nc, int na, int nt, int nu,
                                                       int slab, int nq, float step) {
  const long long entry = static_cast<long long>(blockIdx.x) * kHybridWaves + (threadIdx.x >> 6);
  if (entry >= entries) return;
  const int lane = threadIdx.x & 63;
  const int k = static_cast<int>(entry % nq);
  const int ch = static_cast<int>((entry / nq) % nc);
  const int u = static_cast<int>(entry / nq / nc);
  const float4* b = anchors + (static_cast<size_t>(slab) * nu + u) * na;
  const float shift = static_cast<float>(k) * step;
  for (int t0 = 0; t0 < nt; t0 += 64 * F) {
    const int t = t0 + F * lane;
    const int tt = t < nt ? t : nt - F;
    float2 sum[F];
#pragma unroll
    for (int f = 0; f < F; ++f) sum[f] = make_float2(0.f, 0.f);
    for (int a = 0; a < na; ++a) {
      const float4 g = b[a];
      const float2* trace = raw + (static_cast<size_t>(ch) * na + a) * ns * static_cast<size_t>(nt);
      add_sample<F>(trace, g.x + shift, ns, nt, tt, g.y, g.z, sum);
    }
    store_frames<F>(out + entry * nt, t, nt, sum);
  }
}

// Table row `row` of nq knots at a non-negative position, frames [t, t + F):
// linear between the two knots around it (the last interval extended), or
// Catmull-Rom where the four knots around it exist.
template <int F>
__device__ __forceinline__ void table_value(const float2* table, size_t row, int nq, int nt, int t, float position,
                                            bool cubic, float2 (&out)[F]) {
  const int ip = static_cast<int>(position);
  if (cubic && ip != 0 && ip + 2 < nq) {
    const float2* p = table + (row * nq + ip) * nt + t;
    const float w = position - static_cast<float>(ip), v = 1.f - w;
    float2 p0[F], p1[F], pm[F], p2[F];
    load_frames<F>(p, p0);
    load_frames<F>(p + nt, p1);
    load_frames<F>(p - nt, pm);
    load_frames<F>(p + 2 * nt, p2);
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const float2 value = lerp2(p0[f], p1[f], w);
      const float2 left = add2(add2(pm[f], p1[f]), scale2(p0[f], -2.f));
      const float2 right = add2(add2(p0[f], p2[f]), scale2(p1[f], -2.f));
      out[f] = add2(value, scale2(lerp2(left, right, w), -0.5f * w * v));
    }
    return;
  }
  const int i = min(ip, nq - 2);
  const float2* p = table + (row * nq + i) * nt + t;
  float2 a[F], b[F];
  load_frames<F>(p, a);
  load_frames<F>(p + nt, b);
#pragma unroll
  for (int f = 0; f < F; ++f) out[f] = lerp2(a[f], b[f], position - static_cast<float>(i));
}

// Not-a-knot spline value of plane line `line`, angle a, at knot position pos
// (zero outside the knots), frames [t, t + F): the two-stage planes layout
// (line, knot, angle, frame).
template <int F>
__device__ __forceinline__ void spline_value(const float2* plane, const float2* second, size_t line, int a, int t,
                                             int na, int nk, int nt, float pos, float2 (&out)[F]) {
  const int i = static_cast<int>(floorf(pos));
  if (i < 0 || i >= nk - 1) {
#pragma unroll
    for (int f = 0; f < F; ++f) out[f] = make_float2(0.f, 0.f);
    return;
  }
  const float w = pos - static_cast<float>(i), u = 1.f - w;
  const size_t offset = ((line * nk + i) * na + a) * static_cast<size_t>(nt) + t;
  const size_t stride = static_cast<size_t>(na) * nt;
  float2 p0[F], p1[F], m0[F], m1[F];
  load_frames<F>(plane + offset, p0);
  load_frames<F>(plane + offset + stride, p1);
  load_frames<F>(second + offset, m0);
  load_frames<F>(second + offset + stride, m1);
  const float cu = (u * u * u - u) / 6.f, cw = (w * w * w - w) / 6.f;
#pragma unroll
  for (int f = 0; f < F; ++f)
    out[f] = add2(add2(scale2(p0[f], u), scale2(p1[f], w)), add2(scale2(m0[f], cu), scale2(m1[f], cw)));
}

struct FinishArgs {
  const float2* tables[2];
  const float2* planes[2];
  const float2* second[2];
  const float4* rx[2];
  const float4* tx[2];
  float* power;
  int nx, nz, ny, nc, na, nt, first, depth, nq, nk;
  float inverse_step;
  int cubic;
};

// F frames per lane (F = 2 needs an even frame count): a wave covers 64 F
// frames of its voxel per pass.
template <int F>
__global__ __launch_bounds__(256) void k_hybrid_finish(FinishArgs p) {
  const long long blocks = gridDim.x;
  const long long voxel = xcd_block(blocks) * kHybridWaves + (threadIdx.x >> 6);
  const long long count = static_cast<long long>(p.depth) * p.nx * p.ny;
  if (voxel >= count) return;
  const int lane = threadIdx.x & 63;
  // Depth fastest within the slab, then y, then x.
  const int dz = static_cast<int>(voxel % p.depth);
  const int iy = static_cast<int>((voxel / p.depth) % p.ny);
  const int ix = static_cast<int>(voxel / p.depth / p.ny);
  const int z = p.first + dz;
  const float inv_na = 1.f / static_cast<float>(p.na);
  float total = 0.f;
  for (int t0 = 0; t0 < p.nt; t0 += 64 * F) {
    const int t = t0 + F * lane;
    const int tt = t < p.nt ? t : p.nt - F;
    float2 roots[F], coherent[2][F];
    float magnitude[F], energy[2][F];
#pragma unroll
    for (int f = 0; f < F; ++f) {
      roots[f] = make_float2(0.f, 0.f);
      magnitude[f] = 0.f;
    }
    int active = 0;
#pragma unroll
    for (int conf = 0; conf < 2; ++conf) {
      const int v = conf ? iy : ix, u = conf ? ix : iy;
      const int nv = conf ? p.ny : p.nx, nu = conf ? p.nx : p.ny;
      const float4* rx = p.rx[conf] + (static_cast<size_t>(z) * nv + v) * p.nc;
      const float4* tx = p.tx[conf] + (static_cast<size_t>(z) * nu + u) * p.na;
      for (int ch = 0; ch < p.nc; ++ch) {
        const float4 r = rx[ch];
        if (r.x < 0.f) continue;
        float2 val[F];
        table_value<F>(p.tables[conf], static_cast<size_t>(u) * p.nc + ch, p.nq, p.nt, tt, r.x * p.inverse_step,
                       p.cubic != 0, val);
#pragma unroll
        for (int f = 0; f < F; ++f) {
          const float2 x = rotate(val[f], r.y, r.z);
          const float m = sqrtf(x.x * x.x + x.y * x.y);
          const float w = m > 0.f ? 1.f / sqrtf(m) : 0.f;
          roots[f] = add2(roots[f], scale2(x, w));
          magnitude[f] += m;
        }
        ++active;
      }
#pragma unroll
      for (int f = 0; f < F; ++f) {
        coherent[conf][f] = make_float2(0.f, 0.f);
        energy[conf][f] = 0.f;
      }
      for (int a = 0; a < p.na; ++a) {
        const float4 b = tx[a];
        float2 val[F];
        spline_value<F>(p.planes[conf], p.second[conf], v, a, tt, p.na, p.nk, p.nt, b.x, val);
#pragma unroll
        for (int f = 0; f < F; ++f) {
          const float2 x = rotate(val[f], b.y, b.z);
          coherent[conf][f] = add2(coherent[conf][f], x);
          energy[conf][f] += x.x * x.x + x.y * x.y;
        }
      }
    }
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const float dm = roots[f].x * roots[f].x + roots[f].y * roots[f].y - magnitude[f];
      float power = 0.f;
      if (dm > 0.f && magnitude[f] > 0.f && active > 0) {
        const float2 c0 = coherent[0][f], c1 = coherent[1][f];
        const float r = (c0.x * c0.x + c0.y * c0.y) * inv_na * inv_na;
        const float q = (c1.x * c1.x + c1.y * c1.y) * inv_na * inv_na;
        const float var = fmaxf(0.f, energy[0][f] * inv_na - r) + fmaxf(0.f, energy[1][f] * inv_na - q);
        const float num = sqrtf(r) * sqrtf(q), den = r + q + 0.5f * var - num;
        if (den > 0.f) power = dm * (dm / static_cast<float>(active) / magnitude[f]) * (num / den);
      }
      if (t + f < p.nt) total += power;
    }
  }
  total = wave_sum(total);
  if (lane == 0) p.power[(static_cast<size_t>(ix) * p.nz + z) * p.ny + iy] = total / static_cast<float>(p.nt);
}

bool fits(size_t n) { return n <= 0x7fffffffu; }

// Rows of nt complex64 frames at a and b stay 16-byte aligned.
bool even_frames(size_t nt, const void* a, const void* b) {
  return nt % 2 == 0 && reinterpret_cast<uintptr_t>(a) % 16 == 0 && reinterpret_cast<uintptr_t>(b) % 16 == 0;
}

}  // namespace
}  // namespace rcb

using namespace rcb;

extern "C" rcabeam_status rcabeam_hybrid_planes_device(const void* packed, const float* prx, const float* ptx,
                                                       void* planes, size_t nv, size_t nc, size_t na, size_t ns,
                                                       size_t nt, size_t nk) {
  if (!packed || !prx || !ptx || !planes || ns < 2 || nc < 1 || na < 1 || nt < 1 || nk < 2 || !fits(nc) ||
      !fits(na) || !fits(ns) || !fits(nt) || !fits(nk) || !fits(nv))
    return RCABEAM_ERROR_ARGUMENT;
  const long long entries = static_cast<long long>(nv) * nk * na;
  if (entries == 0) return RCABEAM_SUCCESS;
  const long long blocks = (entries + kHybridWaves - 1) / kHybridWaves;
  if (blocks > 0x7fffffffLL) return RCABEAM_ERROR_ARGUMENT;
  // Two frames per lane where rows keep 16-byte alignment (even nt, aligned buffers).
  auto kernel = even_frames(nt, packed, planes) ? k_hybrid_planes<2> : k_hybrid_planes<1>;
  kernel<<<static_cast<unsigned>(blocks), 256>>>(
      static_cast<const float2*>(packed), reinterpret_cast<const float4*>(prx), ptx, static_cast<float2*>(planes),
      entries, static_cast<int>(nc), static_cast<int>(na), static_cast<int>(ns), static_cast<int>(nt),
      static_cast<int>(nk));
  return launch_status();
}

extern "C" rcabeam_status rcabeam_hybrid_tables_device(const void* packed, const float* anchors, void* tables,
                                                       size_t ns, size_t nc, size_t na, size_t nt, size_t nu,
                                                       size_t slab, size_t nq, float step) {
  if (!packed || !anchors || !tables || ns < 2 || nc < 1 || na < 1 || nt < 1 || nq < 2 || !fits(ns) || !fits(nc) ||
      !fits(na) || !fits(nt) || !fits(nu) || !fits(slab) || !fits(nq) || !std::isfinite(step) || step <= 0.f)
    return RCABEAM_ERROR_ARGUMENT;
  const long long entries = static_cast<long long>(nu) * nc * nq;
  if (entries == 0) return RCABEAM_SUCCESS;
  const long long blocks = (entries + kHybridWaves - 1) / kHybridWaves;
  if (blocks > 0x7fffffffLL) return RCABEAM_ERROR_ARGUMENT;
  auto kernel = even_frames(nt, packed, tables) ? k_hybrid_tables<2> : k_hybrid_tables<1>;
  kernel<<<static_cast<unsigned>(blocks), 256>>>(
      static_cast<const float2*>(packed), reinterpret_cast<const float4*>(anchors), static_cast<float2*>(tables),
      entries, static_cast<int>(ns), static_cast<int>(nc), static_cast<int>(na), static_cast<int>(nt),
      static_cast<int>(nu), static_cast<int>(slab), static_cast<int>(nq), step);
  return launch_status();
}

extern "C" rcabeam_status rcabeam_hybrid_finish_device(const void* tables_rc, const void* tables_cr,
                                                       const void* planes_rc, const void* planes_cr,
                                                       const void* second_rc, const void* second_cr,
                                                       const float* rx_rc, const float* tx_rc, const float* rx_cr,
                                                       const float* tx_cr, float* power, size_t nx, size_t nz,
                                                       size_t ny, size_t nc, size_t na, size_t nt, size_t first,
                                                       size_t depth, size_t nq, size_t nk, float inverse_step,
                                                       int cubic) {
  if (!tables_rc || !tables_cr || !planes_rc || !planes_cr || !second_rc || !second_cr || !rx_rc || !tx_rc ||
      !rx_cr || !tx_cr || !power || nc < 1 || na < 1 || nt < 1 || nq < 2 || nk < 2 || first + depth > nz ||
      !fits(nx) || !fits(nz) || !fits(ny) || !fits(nc) || !fits(na) || !fits(nt) || !fits(nq) || !fits(nk) ||
      !std::isfinite(inverse_step) || inverse_step <= 0.f)
    return RCABEAM_ERROR_ARGUMENT;
  const long long count = static_cast<long long>(depth) * nx * ny;
  if (count == 0) return RCABEAM_SUCCESS;
  long long blocks = (count + kHybridWaves - 1) / kHybridWaves;
  blocks = (blocks + kXcdCount - 1) / kXcdCount * kXcdCount;
  if (blocks > 0x7fffffffLL) return RCABEAM_ERROR_ARGUMENT;
  FinishArgs p;
  p.tables[0] = static_cast<const float2*>(tables_rc);
  p.tables[1] = static_cast<const float2*>(tables_cr);
  p.planes[0] = static_cast<const float2*>(planes_rc);
  p.planes[1] = static_cast<const float2*>(planes_cr);
  p.second[0] = static_cast<const float2*>(second_rc);
  p.second[1] = static_cast<const float2*>(second_cr);
  p.rx[0] = reinterpret_cast<const float4*>(rx_rc);
  p.rx[1] = reinterpret_cast<const float4*>(rx_cr);
  p.tx[0] = reinterpret_cast<const float4*>(tx_rc);
  p.tx[1] = reinterpret_cast<const float4*>(tx_cr);
  p.power = power;
  p.nx = static_cast<int>(nx);
  p.nz = static_cast<int>(nz);
  p.ny = static_cast<int>(ny);
  p.nc = static_cast<int>(nc);
  p.na = static_cast<int>(na);
  p.nt = static_cast<int>(nt);
  p.first = static_cast<int>(first);
  p.depth = static_cast<int>(depth);
  p.nq = static_cast<int>(nq);
  p.nk = static_cast<int>(nk);
  p.inverse_step = inverse_step;
  p.cubic = cubic;
  // Two frames per lane where rows and planes keep 16-byte alignment (even nt).
  if (even_frames(nt, tables_rc, tables_cr) && even_frames(nt, planes_rc, planes_cr) &&
      even_frames(nt, second_rc, second_cr))
    k_hybrid_finish<2><<<static_cast<unsigned>(blocks), 256>>>(p);
  else
    k_hybrid_finish<1><<<static_cast<unsigned>(blocks), 256>>>(p);
  return launch_status();
}
