// Two-stage OPW approximation on MI355X (gfx950): each aperture is beamformed
// onto dense reference planes (lateral line x uniformly spaced reference
// depths, one plane per elevation slab), a uniform not-a-knot cubic spline is
// solved along depth, and every voxel is expanded from the spline after its
// elevation offset is mapped to a depth shift offset * tan(angle / 2).
// Reference: src/rcabeam/two_stage.cu:20-243 (kernels planes / spline / expand
// / finish_stsw) and src/rcabeam/two_stage.py:306-396 (the band loop).
//
// Layouts (complex64 unless stated):
//   raw      (ns, nc, na, nt)                  frames innermost
//   planes   (n_planes * nl, knots, na, chunk) one buffer per aperture
//   IQ out   (nx, nz, ny, nt)
//   St-SW    (10, nx * nz * ny) float32 slow-time sums
//
// Kernels:
//   ts_planes  one wave per (line, angle, block of 8 knots); lanes are frames.
//              The 64 lanes first form the taps of 8 knots x 8 active channels
//              at once (float64 delays split into an exact integer sample, a
//              float32 fraction and a float32 phasor, as the delay-and-sum
//              kernels do); then, channel by channel, the knots are walked in
//              increasing depth.  A channel's sample index never decreases with
//              depth, so the two interpolation rows stay in registers and each
//              knot loads at most one new row (0.67 rows per knot on average at
//              four reference depths per wavelength).
//   ts_spline  one thread per (line, angle, frame): the not-a-knot system's
//              first and last interior rows are diagonal (6 M = rhs), the rest
//              is a Thomas sweep with precomputed factors.
//   ts_expand  one wave per voxel; lanes are frames; per angle the two spline
//              lookups use a voxel position computed in float64 (a float32
//              position at 20 mm carries ~5e-5 knot of rounding).
//   ts_stsw_finish / ts_stsw_normalize  the two-subset St-SW weights from the
//              slow-time sums, normalised by their maximum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "common.hpp"
#include "rcabeam_mi355x.h"

namespace rcb {
namespace {

constexpr int kWaves = 4;           // waves per workgroup
constexpr int kKnotBlock = 8;       // knots per plane wave
constexpr int kChannelGroup = 8;    // channels whose taps one wave forms together (8 x 8 = 64 lanes)
constexpr int kMaxGroups = 4;       // frame groups of 64 per wave (256 frames)
constexpr int kXcds = 8;            // MI355X accelerator complex dies

struct PlaneArgs {
  const float2* raw;
  const float* lateral;   // nl receive-line coordinates
  const float* elements;  // nc receive elements
  const float* angles;
  const float* starts;
  const float* anchors;   // n_planes transmit-axis slab centres
  const float* knots;     // all reference depths
  const unsigned char* same;  // (bands, nl) unchanged-aperture flags
  float2* out;
  const float2* previous;
  int band_id, knot_first, previous_first, previous_count;
  int ns, nc, na, nt, nl, n_planes, npz, chunk, first, count;
  double c, fs, fd;
  float radius;  // aperture radius (< 0: full aperture)
  int baseband;
};

template <int G>
__global__ __launch_bounds__(256) void ts_planes(PlaneArgs a) {
  const int lane = threadIdx.x & 63;
  // A workgroup is (angle, 4 adjacent lines, knot block): adjacent lines read
  // nearly the same sample rows, which the CU's L1 then serves.  Workgroups
  // are dealt to XCDs in contiguous runs (the grid is padded to a multiple of
  // 8), so neighbouring lines also share an L2.
  const int nkb = (a.npz + kKnotBlock - 1) / kKnotBlock;
  const int lines_total = a.n_planes * a.nl;
  const int nlg = (lines_total + kWaves - 1) / kWaves;
  const long long per_xcd = gridDim.x / kXcds;
  const long long lb = static_cast<long long>(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (lb >= static_cast<long long>(a.na) * nlg * nkb) return;
  const int kb = static_cast<int>(lb % nkb);
  const int angle = static_cast<int>(lb / nkb / nlg);
  const int line = static_cast<int>((lb / nkb) % nlg) * kWaves + (threadIdx.x >> 6);
  if (line >= lines_total) return;
  const int il = line % a.nl;
  const int k0 = kb * kKnotBlock;
  const int nk = min(kKnotBlock, a.npz - k0);
  const int f0 = blockIdx.y * 64 * G;
  const size_t plane_row = static_cast<size_t>(a.na) * a.chunk;  // one knot of one line

  // Knots also present in the previous band of a line whose channel set did not
  // change are copied (two_stage.cu:37-45).
  int copy_lo = 0, copy_hi = 0;
  if (a.previous_count > 0 && a.same[static_cast<size_t>(a.band_id) * a.nl + il]) {
    copy_lo = max(k0, a.previous_first - a.knot_first);
    copy_hi = min(k0 + nk, a.previous_first + a.previous_count - a.knot_first);
  }
  for (int k = copy_lo; k < copy_hi; ++k) {
    const size_t dst = (static_cast<size_t>(line) * a.npz + k) * plane_row + static_cast<size_t>(angle) * a.chunk;
    const size_t src = (static_cast<size_t>(line) * a.previous_count + (a.knot_first + k - a.previous_first)) *
                           plane_row + static_cast<size_t>(angle) * a.chunk;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int f = f0 + g * 64 + lane;
      if (f < a.count) a.out[dst + f] = a.previous[src + f];
    }
  }
  unsigned todo = 0;  // knots of this block to beamform
  for (int k = 0; k < nk; ++k)
    if (k0 + k < copy_lo || k0 + k >= copy_hi) todo |= 1u << k;
  if (todo == 0) return;

  // Transmit term of this lane's knot (lane & 7).
  const double c = a.c, fs = a.fs, fd = a.fd;
  const double theta = static_cast<double>(a.angles[angle]);
  const double cos_a = cos(theta), sin_a = sin(theta);
  const double u = static_cast<double>(a.anchors[line / a.nl]);
  const double v = static_cast<double>(a.lateral[il]);
  const int my_k = lane & (kKnotBlock - 1);
  const double zk = static_cast<double>(a.knots[a.knot_first + k0 + min(my_k, nk - 1)]);
  Term tx;
  {
    const double tau = (zk * cos_a + u * sin_a) / c;
    tx = split_delay((tau - static_cast<double>(a.starts[angle])) * fs, a.baseband ? 0.0 : tau * fd);
  }

  float2 acc[kKnotBlock][G];
#pragma unroll
  for (int k = 0; k < kKnotBlock; ++k)
#pragma unroll
    for (int g = 0; g < G; ++g) acc[k][g] = make_float2(0.f, 0.f);

  const size_t row_stride = static_cast<size_t>(a.nc) * a.na * a.nt;
  const float lat32 = a.lateral[il];
  for (int c0 = 0; c0 < a.nc; c0 += 64) {
    const int ch_lane = c0 + lane;
    const bool active = ch_lane < a.nc && (a.radius < 0.f || fabsf(lat32 - a.elements[ch_lane]) <= a.radius);
    uint64_t live = __ballot(active);
    while (live) {
      // Next (up to) 8 active channels of this pass.
      int chans[kChannelGroup];
      int n_in = 0;
#pragma unroll
      for (int j = 0; j < kChannelGroup; ++j) {
        chans[j] = -1;
        if (live) {
          chans[j] = c0 + __builtin_ctzll(live);
          live &= live - 1;
          n_in = j + 1;
        }
      }
      // Lane (cj, k) = (lane >> 3, lane & 7) forms the tap of channel chans[cj] at knot k.
      const int cj = lane >> 3;
      int my_ch = chans[0];
#pragma unroll
      for (int j = 1; j < kChannelGroup; ++j) my_ch = cj == j ? chans[j] : my_ch;
      Tap tap{-1, 0.f, 0.f, 0.f};
      if (my_ch >= 0 && my_k < nk) {
        const double d = v - static_cast<double>(a.elements[my_ch]);
        const double r = sqrt(zk * zk + d * d);
        // Baseband: only the residual (r - z) / c = d^2 / ((r + z) c) of the
        // closest-element carrier is rotated (two_stage.cu:62-68).
        const double cycles = a.baseband ? fd * (d * d) / ((r + zk) * c) : fd * r / c;
        tap = combine(split_delay(r / c * fs, cycles), tx, a.ns);
      }
      for (int j = 0; j < n_in; ++j) {
        const int ch = __builtin_amdgcn_readlane(my_ch, j * kKnotBlock);
        const float2* src = a.raw + (static_cast<size_t>(ch) * a.na + angle) * a.nt + a.first;
        int have = -2;  // row held in lo (hi holds have + 1)
        float2 lo[G], hi[G];
#pragma unroll
        for (int g = 0; g < G; ++g) lo[g] = hi[g] = make_float2(0.f, 0.f);
#pragma unroll
        for (int k = 0; k < kKnotBlock; ++k) {
          if (!((todo >> k) & 1u)) continue;
          const int slot = j * kKnotBlock + k;
          const int i0 = __builtin_amdgcn_readlane(tap.i0, slot);
          if (i0 < 0) continue;
          const float w = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tap.w), slot));
          const float pc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tap.c), slot));
          const float ps = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tap.s), slot));
          if (i0 != have) {
            const int i1 = min(i0 + 1, a.ns - 1);
            if (i0 == have + 1) {
#pragma unroll
              for (int g = 0; g < G; ++g) {
                const int f = f0 + g * 64 + lane;
                lo[g] = hi[g];
                hi[g] = f < a.count ? src[static_cast<size_t>(i1) * row_stride + f] : make_float2(0.f, 0.f);
              }
            } else {
#pragma unroll
              for (int g = 0; g < G; ++g) {
                const int f = f0 + g * 64 + lane;
                lo[g] = f < a.count ? src[static_cast<size_t>(i0) * row_stride + f] : make_float2(0.f, 0.f);
                hi[g] = f < a.count ? src[static_cast<size_t>(i1) * row_stride + f] : make_float2(0.f, 0.f);
              }
            }
            have = i0;
          }
#pragma unroll
          for (int g = 0; g < G; ++g) {
            const float re = fmaf(w, hi[g].x - lo[g].x, lo[g].x);
            const float im = fmaf(w, hi[g].y - lo[g].y, lo[g].y);
            acc[k][g].x = fmaf(-ps, im, fmaf(pc, re, acc[k][g].x));
            acc[k][g].y = fmaf(ps, re, fmaf(pc, im, acc[k][g].y));
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kKnotBlock; ++k) {
    if (!((todo >> k) & 1u)) continue;
    const size_t dst = (static_cast<size_t>(line) * a.npz + k0 + k) * plane_row + static_cast<size_t>(angle) * a.chunk;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int f = f0 + g * 64 + lane;
      if (f < a.count) a.out[dst + f] = acc[k][g];
    }
  }
}

// Not-a-knot second derivatives (times the knot step squared) along knots
// (two_stage.py:119-143, two_stage.cu:86-111).  Row 1 and row n-2 reduce to
// 6 M = rhs once the end conditions M0 = 2 M1 - M2 and M(n-1) = 2 M(n-2) -
// M(n-3) are substituted; rows 2..n-3 are M(i-1) + 4 M(i) + M(i+1) = rhs.
__global__ __launch_bounds__(256) void ts_spline(const float2* __restrict__ p, float2* __restrict__ m,
                                                 const float* __restrict__ upper, const float* __restrict__ inverse,
                                                 long long systems, int npz, int na, int chunk, int count) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= systems) return;
  const long long stride = static_cast<long long>(na) * chunk;
  if (static_cast<int>(t % chunk) >= count) return;
  const long long base = (t / stride) * npz * stride + t % stride;
  float2 pm = p[base], p0 = p[base + stride], prev = make_float2(0.f, 0.f);
  for (int i = 1; i < npz - 1; ++i) {
    const float2 p1 = p[base + (i + 1) * stride];
    float2 rhs = make_float2(6.f * (pm.x - 2.f * p0.x + p1.x), 6.f * (pm.y - 2.f * p0.y + p1.y));
    float2 cur;
    if (i == 1 || i == npz - 2) {
      cur = make_float2(rhs.x / 6.f, rhs.y / 6.f);
    } else {
      const float inv = inverse[i];
      cur = make_float2((rhs.x - prev.x) * inv, (rhs.y - prev.y) * inv);
    }
    m[base + i * stride] = cur;
    prev = cur;
    pm = p0;
    p0 = p1;
  }
  // prev = M(n-2), final.
  for (int i = npz - 3; i >= 1; --i) {
    const float up = upper[i];
    float2 cur = m[base + i * stride];
    cur.x = fmaf(-up, prev.x, cur.x);
    cur.y = fmaf(-up, prev.y, cur.y);
    m[base + i * stride] = cur;
    prev = cur;
  }
  const float2 m1 = m[base + stride], m2 = m[base + 2 * stride];
  m[base] = make_float2(2.f * m1.x - m2.x, 2.f * m1.y - m2.y);
  const float2 mb = m[base + (npz - 2) * stride], mc = m[base + (npz - 3) * stride];
  m[base + (npz - 1) * stride] = make_float2(2.f * mb.x - mc.x, 2.f * mb.y - mc.y);
}

struct ExpandArgs {
  const float2* rc;
  const float2* cr;
  const float2* rc_second;
  const float2* cr_second;
  const float* z;
  const float* slopes;
  void* out;  // IQ (nx, nz, ny, nt) complex64, or St-SW sums (10, nv) float32
  int nx, nz, ny, npz, na, nt, chunk, first, count, depth_start, depth_count, n_planes;
  double plane_z0, plane_dz;
  const int* x_plane;
  const float* x_offset;
  const int* y_plane;
  const float* y_offset;
  const float2* x_phase;  // (nx, na) baseband carrier tables, or null
  const float2* z_phase;  // (nz, na)
  const float2* y_phase;  // (ny, na)
};

struct Lookup {
  long long offset;  // element (line, i, angle, 0) of the plane buffer; < 0: outside the knots
  float ca, cw, sa, sw;
};

__device__ __forceinline__ Lookup make_lookup(long long line, double pos, int angle, int npz, int na, int chunk) {
  Lookup l{-1, 0.f, 0.f, 0.f, 0.f};
  const double fl = floor(pos);
  if (!(fl >= 0.0 && fl < static_cast<double>(npz - 1))) return l;
  const int i = static_cast<int>(fl);
  const double w = pos - fl, am = 1.0 - w;
  l.offset = ((line * npz + i) * na + angle) * static_cast<long long>(chunk);
  l.ca = static_cast<float>(am);
  l.cw = static_cast<float>(w);
  l.sa = static_cast<float>((am * am * am - am) / 6.0);
  l.sw = static_cast<float>((w * w * w - w) / 6.0);
  return l;
}

__device__ __forceinline__ float2 spline_value(const float2* p, const float2* s, const Lookup& l, long long knot_stride,
                                               int f) {
  if (l.offset < 0) return make_float2(0.f, 0.f);
  const float2 p0 = p[l.offset + f], p1 = p[l.offset + knot_stride + f];
  const float2 s0 = s[l.offset + f], s1 = s[l.offset + knot_stride + f];
  float2 r;
  r.x = fmaf(l.ca, p0.x, fmaf(l.cw, p1.x, fmaf(l.sa, s0.x, l.sw * s1.x)));
  r.y = fmaf(l.ca, p0.y, fmaf(l.cw, p1.y, fmaf(l.sa, s0.y, l.sw * s1.y)));
  return r;
}

// Spline values of frames f and f + 1 (f even, rows 16-byte aligned): one
// 16-byte load per row.
__device__ __forceinline__ void spline_pair(const float2* p, const float2* s, const Lookup& l, long long knot_stride,
                                            int f, float2& r0, float2& r1) {
  if (l.offset < 0) {
    r0 = r1 = make_float2(0.f, 0.f);
    return;
  }
  const float4 p0 = *reinterpret_cast<const float4*>(p + l.offset + f);
  const float4 p1 = *reinterpret_cast<const float4*>(p + l.offset + knot_stride + f);
  const float4 s0 = *reinterpret_cast<const float4*>(s + l.offset + f);
  const float4 s1 = *reinterpret_cast<const float4*>(s + l.offset + knot_stride + f);
  r0.x = fmaf(l.ca, p0.x, fmaf(l.cw, p1.x, fmaf(l.sa, s0.x, l.sw * s1.x)));
  r0.y = fmaf(l.ca, p0.y, fmaf(l.cw, p1.y, fmaf(l.sa, s0.y, l.sw * s1.y)));
  r1.x = fmaf(l.ca, p0.z, fmaf(l.cw, p1.z, fmaf(l.sa, s0.z, l.sw * s1.z)));
  r1.y = fmaf(l.ca, p0.w, fmaf(l.cw, p1.w, fmaf(l.sa, s0.w, l.sw * s1.w)));
}

__device__ __forceinline__ float2 coherence(float2 r, float2 q, float var_r, float var_q) {
  const float2 n = cmulc(r, q);
  const float d = cnorm(r) + cnorm(q) + 0.5f * (var_r + var_q) - cabs(n);
  return d > 0.f ? make_float2(n.x / d, n.y / d) : make_float2(0.f, 0.f);
}

// PAIRS (G even, even chunk): frame slot g of a lane is frame
// f0 + (g / 2) 128 + 2 lane + g % 2, so each pair of slots is one 16-byte load
// per plane row; otherwise slot g is frame f0 + 64 g + lane.
template <int G, bool STSW, bool PAIRS>
__global__ __launch_bounds__(256) void ts_expand(ExpandArgs a) {
  const int lane = threadIdx.x & 63;
  const long long task = static_cast<long long>(blockIdx.x) * kWaves + (threadIdx.x >> 6);
  const long long voxels = static_cast<long long>(a.nx) * a.depth_count * a.ny;
  if (task >= voxels) return;
  const int iy = static_cast<int>(task % a.ny);
  const int iz = a.depth_start + static_cast<int>((task / a.ny) % a.depth_count);
  const int ix = static_cast<int>(task / a.ny / a.depth_count);
  const long long rline = ix + (a.n_planes == 1 ? 0LL : static_cast<long long>(a.y_plane[iy]) * a.nx);
  const long long cline = iy + (a.n_planes == 1 ? 0LL : static_cast<long long>(a.x_plane[ix]) * a.ny);
  const double z = static_cast<double>(a.z[iz]);
  const double yo = static_cast<double>(a.y_offset[iy]), xo = static_cast<double>(a.x_offset[ix]);
  const long long knot_stride = static_cast<long long>(a.na) * a.chunk;
  const long long v = (static_cast<long long>(ix) * a.nz + iz) * a.ny + iy;
  const bool baseband = a.z_phase != nullptr;

  float moments[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) moments[i] = 0.f;

  for (int f0 = 0; f0 < a.count; f0 += 64 * G) {
    float2 sum[G];
    float2 rs[2][G], qs[2][G];
    float re[2][G], qe[2][G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      sum[g] = make_float2(0.f, 0.f);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        rs[h][g] = qs[h][g] = make_float2(0.f, 0.f);
        re[h][g] = qe[h][g] = 0.f;
      }
    }
    for (int angle = 0; angle < a.na; ++angle) {
      const double slope = static_cast<double>(a.slopes[angle]);
      const Lookup lr = make_lookup(rline, (z + yo * slope - a.plane_z0) / a.plane_dz, angle, a.npz, a.na, a.chunk);
      const Lookup lq = make_lookup(cline, (z + xo * slope - a.plane_z0) / a.plane_dz, angle, a.npz, a.na, a.chunk);
      float2 ph_r = make_float2(1.f, 0.f), ph_q = make_float2(1.f, 0.f);
      if (baseband) {
        const float2 axial = a.z_phase[static_cast<size_t>(iz) * a.na + angle];
        ph_r = cmul(axial, a.y_phase[static_cast<size_t>(iy) * a.na + angle]);
        ph_q = cmul(axial, a.x_phase[static_cast<size_t>(ix) * a.na + angle]);
      }
      const int h = angle & 1;
      float2 rv[G], qv[G];
      if constexpr (PAIRS) {
#pragma unroll
        for (int g = 0; g < G; g += 2) {
          const int f = f0 + (g / 2) * 128 + 2 * lane;
          if (f >= a.count) continue;
          spline_pair(a.rc, a.rc_second, lr, knot_stride, f, rv[g], rv[g + 1]);
          spline_pair(a.cr, a.cr_second, lq, knot_stride, f, qv[g], qv[g + 1]);
        }
      }
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int f = PAIRS ? f0 + (g / 2) * 128 + 2 * lane + g % 2 : f0 + g * 64 + lane;
        if (f >= a.count) continue;
        float2 r, q;
        if constexpr (PAIRS) {
          r = rv[g];
          q = qv[g];
        } else {
          r = spline_value(a.rc, a.rc_second, lr, knot_stride, f);
          q = spline_value(a.cr, a.cr_second, lq, knot_stride, f);
        }
        if (baseband) {
          r = cmul(r, ph_r);
          q = cmul(q, ph_q);
        }
        if (STSW) {
          // Interleaved subsets: even angles in subset 0, odd in subset 1.
          if (h == 0) {
            rs[0][g] = cadd(rs[0][g], r);
            qs[0][g] = cadd(qs[0][g], q);
            re[0][g] += cnorm(r);
            qe[0][g] += cnorm(q);
          } else {
            rs[1][g] = cadd(rs[1][g], r);
            qs[1][g] = cadd(qs[1][g], q);
            re[1][g] += cnorm(r);
            qe[1][g] += cnorm(q);
          }
        } else {
          sum[g] = cadd(sum[g], cadd(r, q));
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int f = PAIRS ? f0 + (g / 2) * 128 + 2 * lane + g % 2 : f0 + g * 64 + lane;
      if (f >= a.count) continue;
      if (!STSW) {
        static_cast<float2*>(a.out)[v * a.nt + a.first + f] = sum[g];
        continue;
      }
      // Per frame: XDoppler product, per-subset means and variances, and the
      // two disjoint subset pairs (00, 11) and (01, 10) (two_stage.cu:209-229).
      const float2 xd = cmulc(cadd(rs[0][g], rs[1][g]), cadd(qs[0][g], qs[1][g]));
      float2 mr[2], mq[2];
      float vr[2], vq[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float inv = 1.f / static_cast<float>((a.na + 1 - h) / 2);
        mr[h] = make_float2(rs[h][g].x * inv, rs[h][g].y * inv);
        mq[h] = make_float2(qs[h][g].x * inv, qs[h][g].y * inv);
        vr[h] = fmaxf(0.f, re[h][g] * inv - cnorm(mr[h]));
        vq[h] = fmaxf(0.f, qe[h][g] * inv - cnorm(mq[h]));
      }
      moments[0] += xd.x;
      moments[1] += xd.y;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const float2 ca = coherence(mr[0], mq[j], vr[0], vq[j]);
        const float2 cb = coherence(mr[1], mq[1 - j], vr[1], vq[1 - j]);
        const float2 prod = cmulc(ca, cb);
        moments[2 + 4 * j] += prod.x;
        moments[3 + 4 * j] += prod.y;
        moments[4 + 4 * j] += cnorm(ca);
        moments[5 + 4 * j] += cnorm(cb);
      }
    }
  }
  if (STSW) {
    float* stats = static_cast<float*>(a.out);
    const long long nv = static_cast<long long>(a.nx) * a.nz * a.ny;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      const float total = wave_sum(moments[i]);
      if (lane == 0) {
        float* dst = stats + i * nv + v;
        *dst = a.first == 0 ? total : *dst + total;
      }
    }
  }
}

__device__ __forceinline__ float wave_max(float x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmaxf(x, __shfl_xor(x, off));
  return x;
}

__global__ void ts_zero_word(unsigned* word) { *word = 0u; }

// Two-subset St-SW weight and XDoppler power per voxel (two_stage.cu:138-152);
// the largest weight is gathered for the normalisation.
__global__ __launch_bounds__(256) void ts_stsw_finish(const float* __restrict__ stats, float* __restrict__ weights,
                                                      float* __restrict__ power, long long nv, int nt,
                                                      unsigned* __restrict__ peak) {
  const long long v = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  float w = 0.f;
  if (v < nv) {
    float2 acc = make_float2(0.f, 0.f);
    for (int i = 2; i <= 6; i += 4) {
      const float den = sqrtf(stats[(i + 2) * nv + v] * stats[(i + 3) * nv + v]);
      if (den > 0.f) {
        acc.x += stats[i * nv + v] / den;
        acc.y += stats[(i + 1) * nv + v] / den;
      }
    }
    w = hypotf(acc.x, acc.y);
    weights[v] = w;
    power[v] = hypotf(stats[v], stats[nv + v]) / static_cast<float>(nt);
  }
  // Non-negative floats order like their bit patterns.
  const float m = wave_max(w);
  if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax(peak, __float_as_uint(m));
}

__global__ __launch_bounds__(256) void ts_stsw_normalize(float* __restrict__ weights, float* __restrict__ power,
                                                         long long nv, const unsigned* __restrict__ peak) {
  const long long v = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const float p = __uint_as_float(*peak);
  if (!(p > 0.f)) return;
  const float w = weights[v] / p;
  weights[v] = w;
  power[v] *= w;
}

int frame_groups(int count) { return std::min(kMaxGroups, std::max(1, (count + 63) / 64)); }

rcabeam_status launch_planes(const PlaneArgs& a) {
  const int G = frame_groups(a.count);
  const long long nkb = (a.npz + kKnotBlock - 1) / kKnotBlock;
  const long long nlg = (static_cast<long long>(a.n_planes) * a.nl + kWaves - 1) / kWaves;
  const long long blocks = nlg * a.na * nkb;
  if (blocks == 0 || a.count == 0) return RCABEAM_SUCCESS;
  const dim3 grid(static_cast<unsigned>((blocks + kXcds - 1) / kXcds * kXcds),
                  static_cast<unsigned>((a.count + 64 * G - 1) / (64 * G)));
  switch (G) {
    case 1: ts_planes<1><<<grid, 256>>>(a); break;
    case 2: ts_planes<2><<<grid, 256>>>(a); break;
    case 3: ts_planes<3><<<grid, 256>>>(a); break;
    default: ts_planes<4><<<grid, 256>>>(a); break;
  }
  return launch_status();
}

rcabeam_status launch_spline(const float2* p, float2* m, const float* upper, const float* inverse, long long lines,
                             int npz, int na, int chunk, int count) {
  const long long systems = lines * na * chunk;
  if (systems == 0) return RCABEAM_SUCCESS;
  ts_spline<<<static_cast<unsigned>((systems + 255) / 256), 256>>>(p, m, upper, inverse, systems, npz, na, chunk,
                                                                    count);
  return launch_status();
}

rcabeam_status launch_expand(const ExpandArgs& a, bool stsw) {
  const long long voxels = static_cast<long long>(a.nx) * a.depth_count * a.ny;
  if (voxels == 0 || a.count == 0) return RCABEAM_SUCCESS;
  const int G = frame_groups(a.count);
  const unsigned blocks = static_cast<unsigned>((voxels + kWaves - 1) / kWaves);
  // Frame pairs: an even frame-group count, even chunk rows and 16-byte aligned planes.
  auto aligned = [](const void* q) { return reinterpret_cast<uintptr_t>(q) % 16 == 0; };
  const bool pairs = G % 2 == 0 && a.chunk % 2 == 0 && aligned(a.rc) && aligned(a.rc_second) && aligned(a.cr) &&
                     aligned(a.cr_second);
  if (stsw) {
    switch (G) {
      case 1: ts_expand<1, true, false><<<blocks, 256>>>(a); break;
      case 2:
        if (pairs)
          ts_expand<2, true, true><<<blocks, 256>>>(a);
        else
          ts_expand<2, true, false><<<blocks, 256>>>(a);
        break;
      case 3: ts_expand<3, true, false><<<blocks, 256>>>(a); break;
      default:
        if (pairs)
          ts_expand<4, true, true><<<blocks, 256>>>(a);
        else
          ts_expand<4, true, false><<<blocks, 256>>>(a);
        break;
    }
  } else {
    switch (G) {
      case 1: ts_expand<1, false, false><<<blocks, 256>>>(a); break;
      case 2:
        if (pairs)
          ts_expand<2, false, true><<<blocks, 256>>>(a);
        else
          ts_expand<2, false, false><<<blocks, 256>>>(a);
        break;
      case 3: ts_expand<3, false, false><<<blocks, 256>>>(a); break;
      default:
        if (pairs)
          ts_expand<4, false, true><<<blocks, 256>>>(a);
        else
          ts_expand<4, false, false><<<blocks, 256>>>(a);
        break;
    }
  }
  return launch_status();
}

bool fits_int(size_t v) { return v <= static_cast<size_t>(INT32_MAX); }

}  // namespace
}  // namespace rcb

using namespace rcb;

extern "C" rcabeam_status rcabeam_two_stage_planes_device(
    const void* raw, const float* lateral, const float* elements, const float* angles, const float* starts,
    const float* anchors, const float* knots, const unsigned char* same_aperture, void* out, const void* previous,
    int band_id, int knot_first, int previous_first, int previous_count, size_t ns, size_t nc, size_t na, size_t nt,
    size_t nl, size_t n_planes, size_t npz, size_t chunk, size_t first, size_t count, float c, float fs, float fd,
    float aperture_radius, int baseband) {
  if (!raw || !lateral || !elements || !angles || !starts || !anchors || !knots || !out) return RCABEAM_ERROR_ARGUMENT;
  if (previous_count > 0 && (!previous || !same_aperture)) return RCABEAM_ERROR_ARGUMENT;
  if (ns < 2 || nc < 1 || na < 1 || nt < 1 || nl < 1 || n_planes < 1 || npz < 1 || chunk < 1 || count > chunk ||
      first + count > nt || band_id < 0 || knot_first < 0 || previous_count < 0)
    return RCABEAM_ERROR_ARGUMENT;
  if (!fits_int(ns) || !fits_int(nc) || !fits_int(na) || !fits_int(nt) || !fits_int(nl * n_planes) ||
      !fits_int(npz) || !fits_int(chunk))
    return RCABEAM_ERROR_ARGUMENT;
  if (!(c > 0.f) || !(fs > 0.f) || !std::isfinite(fd) || std::isnan(aperture_radius)) return RCABEAM_ERROR_ARGUMENT;
  PlaneArgs a{static_cast<const float2*>(raw), lateral, elements, angles, starts, anchors, knots, same_aperture,
              static_cast<float2*>(out), static_cast<const float2*>(previous), band_id, knot_first, previous_first,
              previous_count, static_cast<int>(ns), static_cast<int>(nc), static_cast<int>(na), static_cast<int>(nt),
              static_cast<int>(nl), static_cast<int>(n_planes), static_cast<int>(npz), static_cast<int>(chunk),
              static_cast<int>(first), static_cast<int>(count), static_cast<double>(c), static_cast<double>(fs),
              static_cast<double>(fd), aperture_radius, baseband ? 1 : 0};
  return launch_planes(a);
}

extern "C" rcabeam_status rcabeam_two_stage_spline_device(const void* planes, void* second, const float* upper,
                                                          const float* inverse, size_t lines, size_t npz, size_t na,
                                                          size_t chunk, size_t count) {
  if (!planes || !second || !upper || !inverse || npz < 4 || count > chunk || !fits_int(npz) || !fits_int(na) ||
      !fits_int(chunk))
    return RCABEAM_ERROR_ARGUMENT;
  return launch_spline(static_cast<const float2*>(planes), static_cast<float2*>(second), upper, inverse,
                       static_cast<long long>(lines), static_cast<int>(npz), static_cast<int>(na),
                       static_cast<int>(chunk), static_cast<int>(count));
}

extern "C" rcabeam_status rcabeam_two_stage_expand_device(
    const void* rc_planes, const void* cr_planes, const void* rc_second, const void* cr_second, const float* z,
    const float* slopes, void* out, int stsw, size_t nx, size_t nz, size_t ny, size_t npz, size_t na, size_t nt,
    size_t chunk, size_t first, size_t count, size_t depth_start, size_t depth_count, float plane_z0, float plane_dz,
    const int* x_plane, const float* x_offset, const int* y_plane, const float* y_offset, size_t n_planes,
    const void* x_phase, const void* z_phase, const void* y_phase) {
  if (!rc_planes || !cr_planes || !rc_second || !cr_second || !z || !slopes || !out || !x_offset || !y_offset)
    return RCABEAM_ERROR_ARGUMENT;
  if (n_planes > 1 && (!x_plane || !y_plane)) return RCABEAM_ERROR_ARGUMENT;
  const bool any_phase = x_phase || z_phase || y_phase;
  if (any_phase && !(x_phase && z_phase && y_phase)) return RCABEAM_ERROR_ARGUMENT;
  if (npz < 2 || count > chunk || first + count > nt || depth_start + depth_count > nz || n_planes < 1 ||
      !(plane_dz > 0.f) || !std::isfinite(plane_z0))
    return RCABEAM_ERROR_ARGUMENT;
  if (!fits_int(nx) || !fits_int(nz) || !fits_int(ny) || !fits_int(npz) || !fits_int(na) || !fits_int(nt) ||
      !fits_int(chunk))
    return RCABEAM_ERROR_ARGUMENT;
  ExpandArgs a{static_cast<const float2*>(rc_planes), static_cast<const float2*>(cr_planes),
               static_cast<const float2*>(rc_second), static_cast<const float2*>(cr_second), z, slopes, out,
               static_cast<int>(nx), static_cast<int>(nz), static_cast<int>(ny), static_cast<int>(npz),
               static_cast<int>(na), static_cast<int>(nt), static_cast<int>(chunk), static_cast<int>(first),
               static_cast<int>(count), static_cast<int>(depth_start), static_cast<int>(depth_count),
               static_cast<int>(n_planes), static_cast<double>(plane_z0), static_cast<double>(plane_dz), x_plane,
               x_offset, y_plane, y_offset, static_cast<const float2*>(x_phase), static_cast<const float2*>(z_phase),
               static_cast<const float2*>(y_phase)};
  return launch_expand(a, stsw != 0);
}

extern "C" rcabeam_status rcabeam_two_stage_stsw_finish_device(const float* stats, float* weights, float* power,
                                                               size_t nv, size_t nt, unsigned* peak) {
  if (!stats || !weights || !power || !peak || nt < 1) return RCABEAM_ERROR_ARGUMENT;
  if (nv == 0) return RCABEAM_SUCCESS;
  const unsigned blocks = static_cast<unsigned>((nv + 255) / 256);
  ts_zero_word<<<1, 1>>>(peak);
  ts_stsw_finish<<<blocks, 256>>>(stats, weights, power, static_cast<long long>(nv), static_cast<int>(nt), peak);
  ts_stsw_normalize<<<blocks, 256>>>(weights, power, static_cast<long long>(nv), peak);
  return launch_status();
}

extern "C" size_t rcabeam_two_stage_buffer_elements(const rcabeam_two_stage_plan* p) {
  if (!p) return 0;
  int most = 0;
  for (size_t b = 0; b < p->n_bands; ++b) most = std::max(most, p->band_knot_count[b]);
  return std::max(p->nx, p->ny) * p->n_planes * static_cast<size_t>(most) * p->na * p->chunk;
}

extern "C" rcabeam_status rcabeam_two_stage_reconstruct_device(const rcabeam_two_stage_plan* p, const void* rc,
                                                               const void* cr, void* buffers, void* out) {
  if (!p || !rc || !cr || !buffers || !out || p->n_bands == 0) return RCABEAM_ERROR_ARGUMENT;
  const size_t per = rcabeam_two_stage_buffer_elements(p);
  float2* buf[4];
  for (int i = 0; i < 4; ++i) buf[i] = static_cast<float2*>(buffers) + i * per;
  // buf[0], buf[1]: RC planes / second derivatives; buf[2], buf[3]: CR.  The
  // roles swap every band so the previous band's planes stay readable while
  // the next band is built (two_stage.py:326-331).
  for (size_t first = 0; first < p->nt; first += p->chunk) {
    const size_t count = std::min(p->chunk, p->nt - first);
    int previous_first = 0, previous_count = 0;
    float2* planes[2] = {buf[1], buf[3]};
    float2* second[2] = {buf[0], buf[2]};
    for (size_t b = 0; b < p->n_bands; ++b) {
      std::swap(planes[0], second[0]);
      std::swap(planes[1], second[1]);
      const int knot_first = p->band_knot_first[b], npz = p->band_knot_count[b];
      for (int config = 0; config < 2; ++config) {
        const size_t nl = config == 0 ? p->nx : p->ny;
        rcabeam_status s = rcabeam_two_stage_planes_device(
            config == 0 ? rc : cr, config == 0 ? p->x : p->y, config == 0 ? p->x_el : p->y_el, p->angles, p->starts,
            config == 0 ? p->y_centers : p->x_centers, p->knots, config == 0 ? p->x_same : p->y_same, planes[config],
            second[config], static_cast<int>(b), knot_first, previous_first, previous_count, p->ns, p->nc, p->na,
            p->nt, nl, p->n_planes, static_cast<size_t>(npz), p->chunk, first, count, p->c, p->fs, p->fd,
            p->band_radius[b], p->baseband);
        if (s != RCABEAM_SUCCESS) return s;
        s = rcabeam_two_stage_spline_device(planes[config], second[config], p->spline_upper, p->spline_inverse,
                                            nl * p->n_planes, static_cast<size_t>(npz), p->na, p->chunk, count);
        if (s != RCABEAM_SUCCESS) return s;
      }
      rcabeam_status s = rcabeam_two_stage_expand_device(
          planes[0], planes[1], second[0], second[1], p->z, p->slopes, out, p->stsw, p->nx, p->nz, p->ny,
          static_cast<size_t>(npz), p->na, p->nt, p->chunk, first, count,
          static_cast<size_t>(p->band_depth_start[b]), static_cast<size_t>(p->band_depth_count[b]),
          p->knots_host[knot_first], p->dz, p->x_plane, p->x_offset, p->y_plane, p->y_offset, p->n_planes,
          p->baseband ? p->x_phase : nullptr, p->baseband ? p->z_phase : nullptr, p->baseband ? p->y_phase : nullptr);
      if (s != RCABEAM_SUCCESS) return s;
      previous_first = knot_first;
      previous_count = npz;
    }
  }
  return RCABEAM_SUCCESS;
}
