// Frame-batched RC/CR delay-and-sum for MI355X (gfx950): OPW, XDoppler, RC-FMAS,
// DMAS-CCF-ACF, St-SW and separate-aperture IQ, behind rcabeam_ensemble_device
// and rcabeam_opw_fp16_device (reference src/rcabeam/ensemble.cu:657-809).
//
// Work decomposition (one 256-thread workgroup = 4 wave64s):
//   * a workgroup owns VB = 4 * VPW consecutive voxels (y fastest, so the voxels
//     of a workgroup share x and z) and ALL frames, so power reductions are
//     complete inside the workgroup: no float atomics, bitwise-deterministic pd;
//   * wave w owns voxels [w*VPW, (w+1)*VPW); lane l owns frames
//     f0 + l + 64*j (j < FPL), so every sample load of a wave reads 64
//     consecutive frames of one (channel, angle, sample) row: fully coalesced in
//     the packed (channels, angles, samples, frames) layout;
//   * receive terms (voxel, channel) and transmit terms (voxel, angle) are built
//     once per workgroup in float64 (common.hpp) and kept in LDS; for each angle
//     the workgroup combines them into a tap table (i0, w, phasor) in LDS, which
//     the waves read as wave-uniform values (scalar registers);
//   * per-angle channel sums are accumulated separately and folded into the
//     total once per angle (two-level summation keeps float32 rounding ~1e-7).
// Blocks are mapped XCD-aware: block b runs on XCD b % 8, so each XCD gets a
// contiguous run of voxel tiles and neighbouring tiles share that XCD's L2.
#include "common.hpp"
#include "ensemble.hpp"
#include "rcabeam_mi355x.h"

#include <type_traits>

namespace rcb {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
// Voxels per wave of the OPW kernel (compile-time tuning knob).
#ifndef RCB_OPW_VPW
#define RCB_OPW_VPW 2
#endif
constexpr size_t kLdsLimit = 64 * 1024;

// Fill rx[vb * nc + ch] and tx[vb * na + a] for one aperture configuration.
__device__ void fill_terms(const EnsembleArgs& p, int cfg, long long v0, int vb_count, Term* rx, Term* tx) {
  const float* el = cfg == 0 ? p.xe : p.ye;
  for (int idx = threadIdx.x; idx < vb_count * p.nc; idx += blockDim.x) {
    int vb = idx / p.nc, ch = idx - vb * p.nc;
    long long v = v0 + vb;
    Term t{kInactive, 0.f, 0.f, 0.f};
    if (v < p.nv) {
      double x = p.scan[3 * v], z = p.scan[3 * v + 1], y = p.scan[3 * v + 2];
      t = receive_term(cfg == 0 ? x : y, z, static_cast<double>(el[ch]), p.c, p.fs, p.fd, p.fn);
    }
    rx[idx] = t;
  }
  for (int idx = threadIdx.x; idx < vb_count * p.na; idx += blockDim.x) {
    int vb = idx / p.na, a = idx - vb * p.na;
    long long v = v0 + vb;
    Term t{0, 0.f, 0.f, 0.f};
    if (v < p.nv) {
      double x = p.scan[3 * v], z = p.scan[3 * v + 1], y = p.scan[3 * v + 2];
      double th = static_cast<double>(p.angles[a]);
      double sn, cs;
      sincos(th, &sn, &cs);
      t = transmit_term(cfg == 0 ? y : x, z, cs, sn, static_cast<double>(p.starts[a]), p.c, p.fs, p.fd);
    }
    tx[idx] = t;
  }
}

// A tap ready for the inner loop: element offset of the first of the two
// sample rows (channel, angle, i0) in the packed layout, interpolation weight
// and phasor.  A tap that contributes nothing has a zero phasor (a real tap's
// phasor has unit modulus) and points at row 0 of its plane with zero weight.
struct RowTap {
  unsigned off;  // ((ch * na + a) * ns + i0) * nt
  float w;
  float c;
  float s;
};

__device__ __forceinline__ RowTap row_tap(const Tap& t, int ch, int na, int a, int ns, int nt) {
  const unsigned plane = (static_cast<unsigned>(ch) * na + a) * static_cast<unsigned>(ns);
  if (t.i0 < 0) return RowTap{plane * static_cast<unsigned>(nt), 0.f, 0.f, 0.f};
  return RowTap{(plane + static_cast<unsigned>(t.i0)) * static_cast<unsigned>(nt), t.w, t.c, t.s};
}

// Accumulate one angle of one aperture into part[VPW][FPL] for the frames of
// this lane.  `taps` holds the VB * nc row taps of the current angle.  All
// 2*VPW*FPL sample loads of a channel are issued before any arithmetic (one
// wait per channel, not per load).  Lanes whose frame is past the end load the
// last frame instead (a valid address) and their sums are never stored.  When
// w == 0 the second row is not needed (this covers i0 == ns - 1, where the
// closed window admits only w == 0) and the first row is read twice.
template <typename T, int VPW, int FPL>
__device__ __forceinline__ void accumulate_angle(const T* __restrict__ data, const RowTap* taps, int nc, int nt,
                                                 int f0, int wave, int lane, float2 (&part)[VPW][FPL]) {
  int tcl[FPL];
#pragma unroll
  for (int j = 0; j < FPL; ++j) {
    const int t = f0 + lane + kWave * j;
    tcl[j] = t < nt ? t : nt - 1;
  }
  for (int ch = 0; ch < nc; ++ch) {
    unsigned off[VPW];
    float w[VPW], pc[VPW], ps[VPW];
    bool any = false;
#pragma unroll
    for (int vv = 0; vv < VPW; ++vv) {
      const RowTap tp = taps[(wave * VPW + vv) * nc + ch];
      off[vv] = static_cast<unsigned>(uniform(static_cast<int>(tp.off)));
      w[vv] = uniform(tp.w);
      pc[vv] = uniform(tp.c);
      ps[vv] = uniform(tp.s);
      any |= (pc[vv] != 0.f) | (ps[vv] != 0.f);
    }
    if (!any) continue;  // every voxel of this wave is outside aperture/window
    float2 A[VPW][FPL], B[VPW][FPL];
#pragma unroll
    for (int vv = 0; vv < VPW; ++vv) {
      const T* base = data + off[vv];
      const T* next = base + (w[vv] != 0.f ? nt : 0);
#pragma unroll
      for (int j = 0; j < FPL; ++j) {
        A[vv][j] = load_sample(base + tcl[j]);
        B[vv][j] = load_sample(next + tcl[j]);
      }
    }
#pragma unroll
    for (int vv = 0; vv < VPW; ++vv) {
#pragma unroll
      for (int j = 0; j < FPL; ++j) {
        const float vx = fmaf(w[vv], B[vv][j].x - A[vv][j].x, A[vv][j].x);
        const float vy = fmaf(w[vv], B[vv][j].y - A[vv][j].y, A[vv][j].y);
        part[vv][j].x = fmaf(pc[vv], vx, fmaf(-ps[vv], vy, part[vv][j].x));
        part[vv][j].y = fmaf(pc[vv], vy, fmaf(ps[vv], vx, part[vv][j].y));
      }
    }
  }
}

// Methods 0 (OPW), 1 (XDoppler), 2 (RC-FMAS), 4 (St-SW), 5 (apertures):
// angle-major traversal (per-angle channel sums are complete before any
// per-angle nonlinearity).
template <int METHOD, typename T, int VPW, int FPL, int K, typename O = float2>
__global__ __launch_bounds__(kBlock) void das_angle_major(EnsembleArgs p, long long tiles) {
  constexpr int VB = kWavesPerBlock * VPW;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  Term* rx = reinterpret_cast<Term*>(lds_raw);
  Term* tx = rx + VB * p.nc;
  RowTap* taps = reinterpret_cast<RowTap*>(tx + VB * p.na);

  const long long tile = xcd_tile(tiles);
  if (tile >= tiles) return;
  const long long v0 = tile * VB;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const T* __restrict__ rc_data = static_cast<const T*>(p.rc);
  const T* __restrict__ cr_data = static_cast<const T*>(p.cr);

  float pw[VPW], pi[VPW];
#pragma unroll
  for (int vv = 0; vv < VPW; ++vv) pw[vv] = pi[vv] = 0.f;

  for (int f0 = 0; f0 < p.nt; f0 += kWave * FPL) {
    float2 r[VPW][FPL], q[VPW][FPL];
    float2 mr[VPW][FPL][K], mq[VPW][FPL][K];
    float vr[VPW][FPL][K], vq[VPW][FPL][K];
#pragma unroll
    for (int vv = 0; vv < VPW; ++vv)
#pragma unroll
      for (int j = 0; j < FPL; ++j) {
        r[vv][j] = q[vv][j] = make_float2(0.f, 0.f);
#pragma unroll
        for (int s = 0; s < K; ++s) {
          mr[vv][j][s] = mq[vv][j][s] = make_float2(0.f, 0.f);
          vr[vv][j][s] = vq[vv][j][s] = 0.f;
        }
      }
    for (int cfg = 0; cfg < 2; ++cfg) {
      __syncthreads();
      fill_terms(p, cfg, v0, VB, rx, tx);
      for (int a = 0; a < p.na; ++a) {
        __syncthreads();
        for (int idx = threadIdx.x; idx < VB * p.nc; idx += blockDim.x) {
          const int vb = idx / p.nc, ch = idx - vb * p.nc;
          taps[idx] = row_tap(combine(rx[idx], tx[vb * p.na + a], p.ns), ch, p.na, a, p.ns, p.nt);
        }
        __syncthreads();
        float2 part[VPW][FPL];
#pragma unroll
        for (int vv = 0; vv < VPW; ++vv)
#pragma unroll
          for (int j = 0; j < FPL; ++j) part[vv][j] = make_float2(0.f, 0.f);
        accumulate_angle<T, VPW, FPL>(cfg == 0 ? rc_data : cr_data, taps, p.nc, p.nt, f0, wave, lane, part);
#pragma unroll
        for (int vv = 0; vv < VPW; ++vv)
#pragma unroll
          for (int j = 0; j < FPL; ++j) {
            float2 s = METHOD == kRcFmas ? sgnsqrt(part[vv][j]) : part[vv][j];
            if (cfg == 0 || METHOD == kOpw)
              r[vv][j] = cadd(r[vv][j], s);
            else
              q[vv][j] = cadd(q[vv][j], s);
            if constexpr (METHOD == kStsw) {
#pragma unroll
              for (int sub = 0; sub < K; ++sub)
                if (a % K == sub) {
                  if (cfg == 0) {
                    mr[vv][j][sub] = cadd(mr[vv][j][sub], s);
                    vr[vv][j][sub] += cnorm(s);
                  } else {
                    mq[vv][j][sub] = cadd(mq[vv][j][sub], s);
                    vq[vv][j][sub] += cnorm(s);
                  }
                }
            }
          }
      }
    }
    // Epilogue for this frame chunk.
#pragma unroll
    for (int vv = 0; vv < VPW; ++vv) {
      const long long v = v0 + wave * VPW + vv;
#pragma unroll
      for (int j = 0; j < FPL; ++j) {
        const int t = f0 + lane + kWave * j;
        if (v >= p.nv || t >= p.nt) continue;
        if constexpr (METHOD == kApertures) {
          p.signal[(2 * v) * p.nt + t] = r[vv][j];
          p.signal[(2 * v + 1) * p.nt + t] = q[vv][j];
        } else if constexpr (METHOD == kOpw) {
          if (p.signal) store_sample(reinterpret_cast<O*>(p.signal) + v * p.nt + t, r[vv][j]);
          pw[vv] += cnorm(r[vv][j]);
        } else if constexpr (METHOD == kRcFmas) {
          const float2 value = cmul(r[vv][j], q[vv][j]);
          if (p.signal) p.signal[v * p.nt + t] = value;
          pw[vv] += cnorm(value);
        } else {  // XDoppler and St-SW
          const float2 value = cmulc(r[vv][j], q[vv][j]);
          if constexpr (METHOD == kXdoppler)
            if (p.signal) p.signal[v * p.nt + t] = value;
          pw[vv] += value.x;
          pi[vv] += value.y;
          if constexpr (METHOD == kStsw) {
            float2 m_r[K], m_q[K];
            float s_r[K], s_q[K];
#pragma unroll
            for (int sub = 0; sub < K; ++sub) {
              const float count = static_cast<float>((p.na - 1 - sub) / K + 1);
              m_r[sub] = make_float2(mr[vv][j][sub].x / count, mr[vv][j][sub].y / count);
              m_q[sub] = make_float2(mq[vv][j][sub].x / count, mq[vv][j][sub].y / count);
              s_r[sub] = fmaxf(vr[vv][j][sub] / count - cnorm(m_r[sub]), 0.f);
              s_q[sub] = fmaxf(vq[vv][j][sub] / count - cnorm(m_q[sub]), 0.f);
            }
#pragma unroll
            for (int i = 0; i < K; ++i)
#pragma unroll
              for (int jj = 0; jj < K; ++jj) {
                const float2 num = cmulc(m_r[i], m_q[jj]);
                const float den = cnorm(m_r[i]) + cnorm(m_q[jj]) + 0.5f * (s_r[i] + s_q[jj]) - cabs(num);
                float2 val = den > 0.f ? make_float2(num.x / den, num.y / den) : make_float2(0.f, 0.f);
                if (!isfinite(val.x) || !isfinite(val.y)) val = make_float2(0.f, 0.f);
                p.workspace[(v * K * K + i * K + jj) * p.nt + t] = val;
              }
          }
        }
      }
    }
  }
  // Power reductions: every lane of a wave holds different frames of the same voxels.
#pragma unroll
  for (int vv = 0; vv < VPW; ++vv) {
    const float sum = wave_sum(pw[vv]);
    const float sum_i = wave_sum(pi[vv]);
    const long long v = v0 + wave * VPW + vv;
    if (lane == 0 && v < p.nv) {
      const float inv = 1.0f / static_cast<float>(p.nt);
      if constexpr (METHOD == kApertures) {
        p.pd[2 * v] = 0.f;
        p.pd[2 * v + 1] = 0.f;
      } else if constexpr (METHOD == kXdoppler || METHOD == kStsw) {
        const float re = sum * inv, im = sum_i * inv;
        p.pd[2 * v] = hypotf(re, im);
        p.pd[2 * v + 1] = im;
      } else {
        p.pd[2 * v] = sum * inv;
        p.pd[2 * v + 1] = 0.f;
      }
    }
  }
}

// Method 3, DMAS-CCF-ACF: channel-major traversal.  Per frame, the angle sums
// of every channel (both apertures) feed the signed-sqrt DMAS and CCF terms;
// per-angle channel sums feed the ACF.  Taps are formed per lane from the
// LDS-resident terms.  Angles beyond MAXA take a second, angle-major pass.
template <typename T, int FPL, int MAXA>
__global__ __launch_bounds__(kBlock) void das_dmas(EnsembleArgs p, long long tiles) {
  constexpr int VB = kWavesPerBlock;  // one voxel per wave
  extern __shared__ __align__(16) unsigned char lds_raw[];
  Term* rx = reinterpret_cast<Term*>(lds_raw);  // [2][VB][nc]
  Term* tx = rx + 2 * VB * p.nc;                // [2][VB][na]

  const long long tile = xcd_tile(tiles);
  if (tile >= tiles) return;
  const long long v0 = tile * VB;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const long long v = v0 + wave;
  const T* data[2] = {static_cast<const T*>(p.rc), static_cast<const T*>(p.cr)};
  fill_terms(p, 0, v0, VB, rx, tx);
  fill_terms(p, 1, v0, VB, rx + VB * p.nc, tx + VB * p.na);
  __syncthreads();
  const Term* rxv[2] = {rx + wave * p.nc, rx + (VB + wave) * p.nc};
  const Term* txv[2] = {tx + wave * p.na, tx + (VB + wave) * p.na};
  float active = 0.f;
  for (int cfg = 0; cfg < 2; ++cfg)
    for (int ch = 0; ch < p.nc; ++ch) active += rxv[cfg][ch].i == kInactive ? 0.f : 1.f;
  const bool single = p.na <= MAXA;
  float total = 0.f;

  auto sample = [&](int cfg, int ch, int a, int t) -> float2 {
    const Tap tp = combine(rxv[cfg][ch], txv[cfg][a], p.ns);
    if (tp.i0 < 0 || t >= p.nt) return make_float2(0.f, 0.f);
    const T* base = data[cfg] + ((static_cast<size_t>(ch) * p.na + a) * p.ns + tp.i0) * p.nt;
    const size_t step = tp.i0 == p.ns - 1 ? 0 : static_cast<size_t>(p.nt);
    const float2 A = load_sample(base + t), B = load_sample(base + step + t);
    const float vx = fmaf(tp.w, B.x - A.x, A.x), vy = fmaf(tp.w, B.y - A.y, A.y);
    return make_float2(tp.c * vx - tp.s * vy, tp.c * vy + tp.s * vx);
  };

  for (int f0 = 0; f0 < p.nt; f0 += kWave * FPL) {
#pragma unroll
    for (int j = 0; j < FPL; ++j) {
      const int t = f0 + lane + kWave * j;
      float2 ang[2][MAXA];
#pragma unroll
      for (int a = 0; a < MAXA; ++a) ang[0][a] = ang[1][a] = make_float2(0.f, 0.f);
      float2 sqrt_sum = make_float2(0.f, 0.f);
      float abs_sum = 0.f;
      for (int ch = 0; ch < p.nc; ++ch) {
        float2 s[2];
        for (int cfg = 0; cfg < 2; ++cfg) {
          float2 acc = make_float2(0.f, 0.f);
          if (rxv[cfg][ch].i != kInactive) {
            if (single) {
#pragma unroll
              for (int a = 0; a < MAXA; ++a)
                if (a < p.na) {
                  const float2 val = sample(cfg, ch, a, t);
                  acc = cadd(acc, val);
                  ang[cfg][a] = cadd(ang[cfg][a], val);
                }
            } else {
              for (int a = 0; a < p.na; ++a) acc = cadd(acc, sample(cfg, ch, a, t));
            }
          }
          s[cfg] = acc;
        }
        sqrt_sum = cadd(sqrt_sum, cadd(sgnsqrt(s[0]), sgnsqrt(s[1])));
        abs_sum += cabs(s[0]) + cabs(s[1]);
      }
      float2 r = make_float2(0.f, 0.f), q = r;
      float er = 0.f, eq = 0.f;
      if (single) {
#pragma unroll
        for (int a = 0; a < MAXA; ++a)
          if (a < p.na) {
            r = cadd(r, ang[0][a]);
            q = cadd(q, ang[1][a]);
            er += cnorm(ang[0][a]);
            eq += cnorm(ang[1][a]);
          }
      } else {
        for (int a = 0; a < p.na; ++a) {
          float2 ar = make_float2(0.f, 0.f), aq = ar;
          for (int ch = 0; ch < p.nc; ++ch) {
            ar = cadd(ar, sample(0, ch, a, t));
            aq = cadd(aq, sample(1, ch, a, t));
          }
          r = cadd(r, ar);
          q = cadd(q, aq);
          er += cnorm(ar);
          eq += cnorm(aq);
        }
      }
      const float dm = cnorm(sqrt_sum) - abs_sum;
      const float wc = (active > 0.f && abs_sum > 0.f) ? fmaxf(dm, 0.f) / active / abs_sum : 0.f;
      const float inv = 1.0f / static_cast<float>(p.na);
      const float nr = cnorm(r) * inv * inv, nq = cnorm(q) * inv * inv;
      const float var = fmaxf(er * inv - nr, 0.f) + fmaxf(eq * inv - nq, 0.f);
      const float num = sqrtf(nr) * sqrtf(nq);
      const float den = nr + nq + 0.5f * var - num;
      float power = den > 0.f ? dm * wc * (num / den) : 0.f;
      if (!isfinite(power)) power = 0.f;
      if (t < p.nt && v < p.nv) {
        if (p.signal) p.signal[v * p.nt + t] = make_float2(power, 0.f);
        total += power;
      }
    }
  }
  total = wave_sum(total);
  if (lane == 0 && v < p.nv) {
    p.pd[2 * v] = total / static_cast<float>(p.nt);
    p.pd[2 * v + 1] = 0.f;
  }
}

// St-SW weights from the K*K normalised-correlation series of every voxel:
// |sum over ordered pairs (i1 < i2 rows, j1 != j2 columns) of
//  <c[i1,j1], c[i2,j2]> / (||c[i1,j1]|| ||c[i2,j2]||)|   (reference src/rcabeam/stsw.py:66-82).
template <int K>
__global__ __launch_bounds__(kBlock) void stsw_finish(const float2* ccf, float* weights, long long nv, int nt) {
  const long long v = static_cast<long long>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (v >= nv) return;
  const float2* base = ccf + v * K * K * nt;
  float inv_norm[K * K];
#pragma unroll
  for (int s = 0; s < K * K; ++s) {
    float e = 0.f;
    for (int t = lane; t < nt; t += kWave) e += cnorm(base[s * nt + t]);
    e = wave_sum(e);
    inv_norm[s] = e > 0.f ? 1.0f / sqrtf(e) : 0.f;
  }
  float2 acc = make_float2(0.f, 0.f);
  for (int t = lane; t < nt; t += kWave) {
    float2 c[K * K];
#pragma unroll
    for (int s = 0; s < K * K; ++s) {
      const float2 raw = base[s * nt + t];
      c[s] = make_float2(raw.x * inv_norm[s], raw.y * inv_norm[s]);
    }
#pragma unroll
    for (int i1 = 0; i1 < K; ++i1)
#pragma unroll
      for (int j1 = 0; j1 < K; ++j1)
#pragma unroll
        for (int i2 = i1 + 1; i2 < K; ++i2)
#pragma unroll
          for (int j2 = 0; j2 < K; ++j2)
            if (j2 != j1) acc = cadd(acc, cmulc(c[i1 * K + j1], c[i2 * K + j2]));
  }
  acc.x = wave_sum(acc.x);
  acc.y = wave_sum(acc.y);
  if (lane == 0) weights[v] = hypotf(acc.x, acc.y);
}

// Number of (voxel, aperture, channel, angle) taps that contribute (inside
// the dynamic aperture and the closed sample window).  Instrumentation for
// algorithmic-flop accounting; one thread per (voxel, aperture, channel).
__global__ void k_count_taps(EnsembleArgs p, unsigned long long* out) {
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned long long n = 0;
  if (idx < p.nv * 2 * p.nc) {
    const long long v = idx / (2 * p.nc);
    const int cfg = static_cast<int>((idx / p.nc) % 2);
    const int ch = static_cast<int>(idx % p.nc);
    const double x = p.scan[3 * v], z = p.scan[3 * v + 1], y = p.scan[3 * v + 2];
    const float* el = cfg == 0 ? p.xe : p.ye;
    const Term r = receive_term(cfg == 0 ? x : y, z, static_cast<double>(el[ch]), p.c, p.fs, p.fd, p.fn);
    if (r.i != kInactive) {
      for (int a = 0; a < p.na; ++a) {
        double sn, cs;
        sincos(static_cast<double>(p.angles[a]), &sn, &cs);
        const Term t = transmit_term(cfg == 0 ? y : x, z, cs, sn, static_cast<double>(p.starts[a]), p.c, p.fs, p.fd);
        n += combine(r, t, p.ns).i0 >= 0 ? 1 : 0;
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0 && n) atomicAdd(out, n);
}

// ---- host dispatch ---------------------------------------------------------

template <int METHOD, typename T, int VPW, int K, typename O = float2>
static hipError_t launch_angle_major(const EnsembleArgs& p) {
  constexpr int VB = kWavesPerBlock * VPW;
  // Row offsets are 32-bit element indices into one packed aperture.
  if (static_cast<unsigned long long>(p.ns) * p.nc * p.na * p.nt > 0xffffffffULL) return hipErrorInvalidValue;
  const size_t lds = static_cast<size_t>(VB) * (2 * p.nc + p.na) * 16;
  if (lds > kLdsLimit) return hipErrorInvalidValue;
  const long long tiles = (p.nv + VB - 1) / VB;
  const long long grid = ((tiles + 7) / 8) * 8;
  if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
  const int fpl = p.nt <= kWave ? 1 : (p.nt <= 2 * kWave ? 2 : 4);
  switch (fpl) {
    case 1:
      das_angle_major<METHOD, T, VPW, 1, K, O><<<dim3(static_cast<unsigned>(grid)), kBlock, lds>>>(p, tiles);
      break;
    case 2:
      das_angle_major<METHOD, T, VPW, 2, K, O><<<dim3(static_cast<unsigned>(grid)), kBlock, lds>>>(p, tiles);
      break;
    default:
      das_angle_major<METHOD, T, VPW, 4, K, O><<<dim3(static_cast<unsigned>(grid)), kBlock, lds>>>(p, tiles);
      break;
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_dmas(const EnsembleArgs& p) {
  constexpr int VB = kWavesPerBlock;
  const size_t lds = static_cast<size_t>(2 * VB) * (p.nc + p.na) * 16;
  if (lds > kLdsLimit) return hipErrorInvalidValue;
  const long long tiles = (p.nv + VB - 1) / VB;
  const long long grid = ((tiles + 7) / 8) * 8;
  if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
  das_dmas<T, 1, 16><<<dim3(static_cast<unsigned>(grid)), kBlock, lds>>>(p, tiles);
  return hipGetLastError();
}

template <typename T>
constexpr int storage_code() {
  return std::is_same<T, float2>::value ? RCABEAM_STORAGE_C64
                                        : (std::is_same<T, __half2>::value ? RCABEAM_STORAGE_FP16 : RCABEAM_STORAGE_BF16);
}

// Float32-accurate coefficients for a channel storage: FP16 samples use dot2
// with hi + lo coefficients, the others widen samples to float32.
static int accurate_coefficients(int storage) { return storage == RCABEAM_STORAGE_FP16 ? kCoefSplit : kCoefFloat; }

template <typename T>
static hipError_t dispatch(const EnsembleArgs& p, int method, int k) {
  if (method == kOpw || method == kApertures) {
    const hipError_t e = launch_opw_tile(p, storage_code<T>(), accurate_coefficients(storage_code<T>()),
                                         RCABEAM_STORAGE_C64, method == kApertures);
    if (e != hipErrorNotSupported) return e;
  }
  switch (method) {
    case kOpw:
      return launch_angle_major<kOpw, T, RCB_OPW_VPW, 1>(p);
    case kXdoppler:
      return launch_angle_major<kXdoppler, T, 2, 1>(p);
    case kRcFmas:
      return launch_angle_major<kRcFmas, T, 2, 1>(p);
    case kApertures:
      return launch_angle_major<kApertures, T, 2, 1>(p);
    case kDmas:
      return launch_dmas<T>(p);
    case kStsw: {
      hipError_t e;
      const long long blocks = (p.nv + kWavesPerBlock - 1) / kWavesPerBlock;
      if (k == 2) {
        e = launch_angle_major<kStsw, T, 1, 2>(p);
        if (e == hipSuccess) stsw_finish<2><<<dim3(static_cast<unsigned>(blocks)), kBlock>>>(p.workspace, p.weights, p.nv, p.nt);
      } else if (k == 3) {
        e = launch_angle_major<kStsw, T, 1, 3>(p);
        if (e == hipSuccess) stsw_finish<3><<<dim3(static_cast<unsigned>(blocks)), kBlock>>>(p.workspace, p.weights, p.nv, p.nt);
      } else {
        e = launch_angle_major<kStsw, T, 1, 4>(p);
        if (e == hipSuccess) stsw_finish<4><<<dim3(static_cast<unsigned>(blocks)), kBlock>>>(p.workspace, p.weights, p.nv, p.nt);
      }
      if (e != hipSuccess) return e;
      return hipGetLastError();
    }
    default:
      return hipErrorInvalidValue;
  }
}

static bool finite_scalar(float x) { return x == x && x - x == 0.0f; }

// Shared validation of reference src/rcabeam/ensemble.cu:664-675.
static bool valid_ensemble(const void* rc, const void* cr, const float* scan, const float* xe, const float* ye,
                           const float* angles, const float* starts, const float* pd, size_t ns, size_t nc,
                           size_t na, size_t nt, size_t nv, float c, float fs, float fd, float fn) {
  if (!rc || !cr || !scan || !xe || !ye || !angles || !starts || !pd) return false;
  if (ns < 2 || !nc || !na || !nt || !nv) return false;
  if (!(c > 0) || !(fs > 0) || !finite_scalar(c) || !finite_scalar(fs) || !finite_scalar(fd) || !finite_scalar(fn))
    return false;
  // Index arithmetic in the kernels keeps per-axis sizes in int.
  if (ns > 0x7fffffff || nc > 0x7fffffff || na > 0x7fffffff || nt > 0x7fffffff) return false;
  return true;
}

static EnsembleArgs make_args(const void* rc, const void* cr, const float* scan, const float* xe, const float* ye,
                              const float* angles, const float* starts, float* pd, void* signal, float* weights,
                              void* workspace, size_t ns, size_t nc, size_t na, size_t nt, size_t nv, float c,
                              float fs, float fd, float fn) {
  EnsembleArgs p;
  p.rc = rc;
  p.cr = cr;
  p.scan = scan;
  p.xe = xe;
  p.ye = ye;
  p.angles = angles;
  p.starts = starts;
  p.pd = pd;
  p.signal = static_cast<float2*>(signal);
  p.weights = weights;
  p.workspace = static_cast<float2*>(workspace);
  p.ns = static_cast<int>(ns);
  p.nc = static_cast<int>(nc);
  p.na = static_cast<int>(na);
  p.nt = static_cast<int>(nt);
  p.nv = static_cast<long long>(nv);
  p.c = c;
  p.fs = fs;
  p.fd = fd;
  p.fn = fn;
  return p;
}

static rcabeam_status to_status(hipError_t e) {
  if (e == hipSuccess) return RCABEAM_SUCCESS;
  if (e == hipErrorInvalidValue) return RCABEAM_ERROR_ARGUMENT;
  return RCABEAM_ERROR_CUDA;
}

}  // namespace rcb

using namespace rcb;

extern "C" rcabeam_status rcabeam_ensemble_device(const void* rc, const void* cr, const float* scan, const float* xe,
                                                  const float* ye, const float* angles, const float* starts,
                                                  float* pd, void* signal, float* weights, void* workspace,
                                                  void* geometry_workspace, size_t ns, size_t nc, size_t na,
                                                  size_t nt, size_t nv, float c, float fs, float fd, float fn,
                                                  int method, int k) {
  if (!valid_ensemble(rc, cr, scan, xe, ye, angles, starts, pd, ns, nc, na, nt, nv, c, fs, fd, fn))
    return RCABEAM_ERROR_ARGUMENT;
  if (method < 0 || method > 5 || (method == kApertures && !signal) || !geometry_workspace)
    return RCABEAM_ERROR_ARGUMENT;
  if (method == kStsw && (!weights || !workspace || signal || k < 2 || k > 4 || na < static_cast<size_t>(k)))
    return RCABEAM_ERROR_ARGUMENT;
  EnsembleArgs p = make_args(rc, cr, scan, xe, ye, angles, starts, pd, signal, weights, workspace, ns, nc, na, nt,
                             nv, c, fs, fd, fn);
  return to_status(dispatch<float2>(p, method, k));
}

extern "C" rcabeam_status rcabeam_opw_fp16_device(const void* rc, const void* cr, const float* scan, const float* xe,
                                                  const float* ye, const float* angles, const float* starts,
                                                  float* pd, void* signal, void* geometry_workspace, size_t ns,
                                                  size_t nc, size_t na, size_t nt, size_t nv, float c, float fs,
                                                  float fd, float fn) {
  if (reinterpret_cast<uintptr_t>(rc) % 4 || reinterpret_cast<uintptr_t>(cr) % 4) return RCABEAM_ERROR_ARGUMENT;
  if (!valid_ensemble(rc, cr, scan, xe, ye, angles, starts, pd, ns, nc, na, nt, nv, c, fs, fd, fn) ||
      !geometry_workspace)
    return RCABEAM_ERROR_ARGUMENT;
  EnsembleArgs p = make_args(rc, cr, scan, xe, ye, angles, starts, pd, signal, nullptr, nullptr, ns, nc, na, nt, nv,
                             c, fs, fd, fn);
  return to_status(dispatch<__half2>(p, kOpw, 2));
}

extern "C" rcabeam_status rcabeam_opw_bf16_device(const void* rc, const void* cr, const float* scan, const float* xe,
                                                  const float* ye, const float* angles, const float* starts,
                                                  float* pd, void* signal, size_t ns, size_t nc, size_t na,
                                                  size_t nt, size_t nv, float c, float fs, float fd, float fn) {
  if (reinterpret_cast<uintptr_t>(rc) % 4 || reinterpret_cast<uintptr_t>(cr) % 4) return RCABEAM_ERROR_ARGUMENT;
  if (!valid_ensemble(rc, cr, scan, xe, ye, angles, starts, pd, ns, nc, na, nt, nv, c, fs, fd, fn))
    return RCABEAM_ERROR_ARGUMENT;
  EnsembleArgs p = make_args(rc, cr, scan, xe, ye, angles, starts, pd, signal, nullptr, nullptr, ns, nc, na, nt, nv,
                             c, fs, fd, fn);
  return to_status(dispatch<bf16x2>(p, kOpw, 2));
}

extern "C" rcabeam_status rcabeam_count_active_taps_device(const float* scan, const float* xe, const float* ye,
                                                           const float* angles, const float* starts, size_t ns,
                                                           size_t nc, size_t na, size_t nv, float c, float fs,
                                                           float fd, float fn, unsigned long long* out) {
  if (!scan || !xe || !ye || !angles || !starts || !out || ns < 2 || !nc || !na || !nv) return RCABEAM_ERROR_ARGUMENT;
  EnsembleArgs p = make_args(nullptr, nullptr, scan, xe, ye, angles, starts, nullptr, nullptr, nullptr, nullptr, ns,
                             nc, na, 1, nv, c, fs, fd, fn);
  const long long n = static_cast<long long>(nv) * 2 * static_cast<long long>(nc);
  k_count_taps<<<dim3(static_cast<unsigned>((n + 255) / 256)), 256>>>(p, out);
  return launch_status();
}

template <typename T>
static hipError_t dispatch_opw_out(const EnsembleArgs& p, int out_storage) {
  switch (out_storage) {
    case RCABEAM_STORAGE_C64:
      return launch_angle_major<kOpw, T, RCB_OPW_VPW, 1, float2>(p);
    case RCABEAM_STORAGE_FP16:
      return launch_angle_major<kOpw, T, RCB_OPW_VPW, 1, __half2>(p);
    case RCABEAM_STORAGE_BF16:
      return launch_angle_major<kOpw, T, RCB_OPW_VPW, 1, bf16x2>(p);
    default:
      return hipErrorInvalidValue;
  }
}

extern "C" rcabeam_status rcabeam_opw_mixed_device(const void* rc, const void* cr, const float* scan, const float* xe,
                                                   const float* ye, const float* angles, const float* starts,
                                                   float* pd, void* signal, int channel_storage, int signal_storage,
                                                   int coefficients, size_t ns, size_t nc, size_t na, size_t nt,
                                                   size_t nv, float c, float fs, float fd, float fn) {
  if (reinterpret_cast<uintptr_t>(rc) % 4 || reinterpret_cast<uintptr_t>(cr) % 4) return RCABEAM_ERROR_ARGUMENT;
  if (!valid_ensemble(rc, cr, scan, xe, ye, angles, starts, pd, ns, nc, na, nt, nv, c, fs, fd, fn))
    return RCABEAM_ERROR_ARGUMENT;
  if (channel_storage < RCABEAM_STORAGE_C64 || channel_storage > RCABEAM_STORAGE_BF16 ||
      signal_storage < RCABEAM_STORAGE_C64 || signal_storage > RCABEAM_STORAGE_BF16 ||
      (coefficients != RCABEAM_COEF_FLOAT32 && coefficients != RCABEAM_COEF_16BIT))
    return RCABEAM_ERROR_ARGUMENT;
  EnsembleArgs p = make_args(rc, cr, scan, xe, ye, angles, starts, pd, signal, nullptr, nullptr, ns, nc, na, nt, nv,
                             c, fs, fd, fn);
  const int coef = coefficients == RCABEAM_COEF_16BIT && channel_storage != RCABEAM_STORAGE_C64
                       ? kCoefSingle
                       : accurate_coefficients(channel_storage);
  hipError_t e = launch_opw_tile(p, channel_storage, coef, signal_storage, false);
  if (e != hipErrorNotSupported) return to_status(e);
  switch (channel_storage) {
    case RCABEAM_STORAGE_C64:
      e = dispatch_opw_out<float2>(p, signal_storage);
      break;
    case RCABEAM_STORAGE_FP16:
      e = dispatch_opw_out<__half2>(p, signal_storage);
      break;
    case RCABEAM_STORAGE_BF16:
      e = dispatch_opw_out<bf16x2>(p, signal_storage);
      break;
    default:
      e = hipErrorInvalidValue;
  }
  return to_status(e);
}

extern "C" rcabeam_status rcabeam_opw_grid_device(const void* rc, const void* cr, int channel_storage, int coefficients,
                                                  const float* x, const float* z, const float* y, size_t nx,
                                                  size_t nz, size_t ny, const float* xe, const float* ye,
                                                  const float* angles, const float* starts, float* pd, void* signal,
                                                  int signal_storage, int method, size_t ns, size_t nc, size_t na,
                                                  size_t nt, float c, float fs, float fd, float fn) {
  if (!x || !z || !y || !nx || !nz || !ny || nx > 0x7fffffff || nz > 0x7fffffff || ny > 0x7fffffff)
    return RCABEAM_ERROR_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(rc) % 4 || reinterpret_cast<uintptr_t>(cr) % 4) return RCABEAM_ERROR_ARGUMENT;
  const size_t nv = nx * nz * ny;
  // The scan list is not read on a grid; any non-null pointer passes validation.
  if (!valid_ensemble(rc, cr, x, xe, ye, angles, starts, pd, ns, nc, na, nt, nv, c, fs, fd, fn))
    return RCABEAM_ERROR_ARGUMENT;
  if ((method != kOpw && method != kApertures) || (method == kApertures && (!signal || channel_storage != 0)) ||
      (coefficients != RCABEAM_COEF_FLOAT32 && coefficients != RCABEAM_COEF_16BIT) || channel_storage < 0 ||
      channel_storage > 2 || signal_storage < 0 || signal_storage > 2)
    return RCABEAM_ERROR_ARGUMENT;
  EnsembleArgs p = make_args(rc, cr, nullptr, xe, ye, angles, starts, pd, signal, nullptr, nullptr, ns, nc, na, nt, nv,
                             c, fs, fd, fn);
  const int coef = coefficients == RCABEAM_COEF_16BIT && channel_storage != RCABEAM_STORAGE_C64
                       ? kCoefSingle
                       : accurate_coefficients(channel_storage);
  hipError_t e = hipErrorNotSupported;
  // OPW: 16-bit channels on MFMA; OPW and separate apertures of complex64
  // channels on the VALU grid kernel; BF16 channels with float32-accurate
  // coefficients and shapes those kernels do not take: the VALU tile kernel.
  if (method == kOpw && channel_storage != RCABEAM_STORAGE_C64)
    e = launch_opw_grid_mfma(p, x, z, y, static_cast<int>(nx), static_cast<int>(nz), static_cast<int>(ny),
                             channel_storage, coef, signal_storage);
  if (channel_storage == RCABEAM_STORAGE_C64)
    e = launch_opw_grid_valu(p, x, z, y, static_cast<int>(nx), static_cast<int>(nz), static_cast<int>(ny),
                             signal_storage, method == kApertures);
  if (e == hipErrorNotSupported)
    e = launch_opw_grid(p, x, z, y, static_cast<int>(nx), static_cast<int>(nz), static_cast<int>(ny),
                        channel_storage, coef, signal_storage, method == kApertures);
  if (e == hipErrorNotSupported) return RCABEAM_ERROR_ARGUMENT;
  return to_status(e);
}

extern "C" rcabeam_status rcabeam_opw_grid_split_device(const void* rc, const void* cr, const unsigned* absmax,
                                                        const float* x, const float* z, const float* y, size_t nx,
                                                        size_t nz, size_t ny, const float* xe, const float* ye,
                                                        const float* angles, const float* starts, float* pd,
                                                        void* signal, int signal_storage, size_t ns, size_t nc,
                                                        size_t na, size_t nt, float c, float fs, float fd, float fn) {
  if (!x || !z || !y || !nx || !nz || !ny || nx > 0x7fffffff || nz > 0x7fffffff || ny > 0x7fffffff || !absmax)
    return RCABEAM_ERROR_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(rc) % 8 || reinterpret_cast<uintptr_t>(cr) % 8 ||
      reinterpret_cast<uintptr_t>(absmax) % 4)
    return RCABEAM_ERROR_ARGUMENT;
  const size_t nv = nx * nz * ny;
  if (!valid_ensemble(rc, cr, x, xe, ye, angles, starts, pd, ns, nc, na, nt, nv, c, fs, fd, fn) ||
      signal_storage < RCABEAM_STORAGE_C64 || signal_storage > RCABEAM_STORAGE_BF16)
    return RCABEAM_ERROR_ARGUMENT;
  EnsembleArgs p = make_args(rc, cr, nullptr, xe, ye, angles, starts, pd, signal, nullptr, nullptr, ns, nc, na, nt, nv,
                             c, fs, fd, fn);
  const hipError_t e = launch_opw_grid_mfma(p, x, z, y, static_cast<int>(nx), static_cast<int>(nz),
                                            static_cast<int>(ny), RCABEAM_STORAGE_FP16_SPLIT, kCoefSplit,
                                            signal_storage, absmax);
  if (e == hipErrorNotSupported) return RCABEAM_ERROR_ARGUMENT;
  return to_status(e);
}

extern "C" rcabeam_status rcabeam_opw_grid_aperture_device(const void* rc, const void* cr, const float* x,
                                                           const float* z, const float* y, size_t nx, size_t nz,
                                                           size_t ny, const float* xe, const float* ye,
                                                           const float* angles, const float* starts, float* pd,
                                                           void* signal, int aperture, int accumulate, size_t ns,
                                                           size_t nc, size_t na, size_t nt, float c, float fs,
                                                           float fd, float fn) {
  if (!x || !z || !y || !nx || !nz || !ny || nx > 0x7fffffff || nz > 0x7fffffff || ny > 0x7fffffff || !signal ||
      (aperture != 0 && aperture != 1))
    return RCABEAM_ERROR_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(rc) % 8 || reinterpret_cast<uintptr_t>(cr) % 8 ||
      reinterpret_cast<uintptr_t>(signal) % 8)
    return RCABEAM_ERROR_ARGUMENT;
  const size_t nv = nx * nz * ny;
  if (!valid_ensemble(rc, cr, x, xe, ye, angles, starts, pd, ns, nc, na, nt, nv, c, fs, fd, fn))
    return RCABEAM_ERROR_ARGUMENT;
  EnsembleArgs p = make_args(rc, cr, nullptr, xe, ye, angles, starts, pd, signal, nullptr, nullptr, ns, nc, na, nt, nv,
                             c, fs, fd, fn);
  const int q0 = aperture == 0 ? 0 : static_cast<int>(nc);
  const hipError_t e = launch_opw_grid_valu(p, x, z, y, static_cast<int>(nx), static_cast<int>(nz),
                                            static_cast<int>(ny), RCABEAM_STORAGE_C64, false, q0,
                                            q0 + static_cast<int>(nc), accumulate != 0);
  if (e == hipErrorNotSupported) return RCABEAM_ERROR_ARGUMENT;
  return to_status(e);
}

namespace rcb {
namespace {

// XDoppler from separate-aperture IQ (voxels, 2, frames): S_t = RC_t conj(CR_t),
// PD = |mean_t S| with the imaginary mean in the odd entry, as the XDoppler
// epilogue of das_angle_major forms them.  One wave per voxel.
__global__ __launch_bounds__(256) void k_xdoppler_apertures(const float2* __restrict__ ap, long long nv, int nt,
                                                            float* __restrict__ pd, float2* __restrict__ signal) {
  const long long v = static_cast<long long>(blockIdx.x) * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (v >= nv) return;
  const float2* rc = ap + 2 * v * nt;
  const float2* cr = rc + nt;
  float re = 0.f, im = 0.f;
  for (int t = lane; t < nt; t += kWave) {
    const float2 value = cmulc(rc[t], cr[t]);
    if (signal) signal[v * nt + t] = value;
    re += value.x;
    im += value.y;
  }
  re = wave_sum(re);
  im = wave_sum(im);
  if (lane == 0) {
    const float inv = 1.0f / static_cast<float>(nt);
    pd[2 * v] = hypotf(re * inv, im * inv);
    pd[2 * v + 1] = im * inv;
  }
}

}  // namespace
}  // namespace rcb

extern "C" rcabeam_status rcabeam_xdoppler_from_apertures_device(const void* apertures, size_t nv, size_t nt,
                                                                 float* pd, void* signal) {
  if (!apertures || !pd || !nv || !nt || nt > 0x7fffffff || nv > (1ull << 40)) return RCABEAM_ERROR_ARGUMENT;
  const unsigned grid = static_cast<unsigned>((nv + 3) / 4);
  k_xdoppler_apertures<<<grid, 256>>>(static_cast<const float2*>(apertures), static_cast<long long>(nv),
                                      static_cast<int>(nt), pd, static_cast<float2*>(signal));
  return launch_status();
}

extern "C" rcabeam_status rcabeam_opw_adjoint_grid_device(const void* rc, const void* cr, const float* x,
                                                          const float* z, const float* y, size_t nx, size_t nz,
                                                          size_t ny, const float* xe, const float* ye,
                                                          const float* angles, const float* starts,
                                                          const void* grad_iq, void* grad_rc, void* grad_cr,
                                                          double* grad_c, size_t ns, size_t nc, size_t na, size_t nt,
                                                          float c, float fs, float fd, float fn) {
  if (!x || !z || !y || !nx || !nz || !ny || nx > 0x7fffffff || nz > 0x7fffffff || ny > 0x7fffffff)
    return RCABEAM_ERROR_ARGUMENT;
  if (!grad_iq || !grad_rc || !grad_cr || !grad_c) return RCABEAM_ERROR_ARGUMENT;
  const size_t nv = nx * nz * ny;
  float dummy_pd = 0.f;  // validation only: pd is not written by the adjoint
  if (!valid_ensemble(rc, cr, x, xe, ye, angles, starts, &dummy_pd, ns, nc, na, nt, nv, c, fs, fd, fn))
    return RCABEAM_ERROR_ARGUMENT;
  EnsembleArgs p = make_args(rc, cr, nullptr, xe, ye, angles, starts, nullptr, nullptr, nullptr, nullptr, ns, nc, na,
                             nt, nv, c, fs, fd, fn);
  const hipError_t e = launch_opw_adjoint_grid(p, x, z, y, static_cast<int>(nx), static_cast<int>(nz),
                                               static_cast<int>(ny), static_cast<const float2*>(grad_iq),
                                               static_cast<float2*>(grad_rc), static_cast<float2*>(grad_cr), grad_c);
  if (e == hipErrorNotSupported) return RCABEAM_ERROR_ARGUMENT;
  return to_status(e);
}

extern "C" size_t rcabeam_opw_adjoint_workspace_bytes(size_t nx, size_t nz, size_t ny, size_t ns, size_t nc,
                                                      size_t na, size_t nt) {
  if (!nx || !nz || !ny || !ns || !nc || !na || !nt || nx > 0x7fffffff || nz > 0x7fffffff || ny > 0x7fffffff)
    return 0;
  EnsembleArgs p{};
  p.ns = static_cast<int>(ns);
  p.nc = static_cast<int>(nc);
  p.na = static_cast<int>(na);
  p.nt = static_cast<int>(nt);
  return adjoint_workspace_bytes(p, static_cast<int>(nx), static_cast<int>(nz), static_cast<int>(ny));
}

extern "C" rcabeam_status rcabeam_opw_adjoint_grid_deterministic_device(
    const void* rc, const void* cr, const float* x, const float* z, const float* y, size_t nx, size_t nz, size_t ny,
    const float* xe, const float* ye, const float* angles, const float* starts, const void* grad_iq, void* grad_rc,
    void* grad_cr, double* grad_c, void* workspace, size_t workspace_bytes, size_t ns, size_t nc, size_t na,
    size_t nt, float c, float fs, float fd, float fn) {
  if (!x || !z || !y || !nx || !nz || !ny || nx > 0x7fffffff || nz > 0x7fffffff || ny > 0x7fffffff)
    return RCABEAM_ERROR_ARGUMENT;
  if (!grad_iq || !grad_rc || !grad_cr || !grad_c || !workspace || reinterpret_cast<uintptr_t>(workspace) % 256)
    return RCABEAM_ERROR_ARGUMENT;
  const size_t nv = nx * nz * ny;
  float dummy_pd = 0.f;  // validation only: pd is not written by the adjoint
  if (!valid_ensemble(rc, cr, x, xe, ye, angles, starts, &dummy_pd, ns, nc, na, nt, nv, c, fs, fd, fn))
    return RCABEAM_ERROR_ARGUMENT;
  if (workspace_bytes < rcabeam_opw_adjoint_workspace_bytes(nx, nz, ny, ns, nc, na, nt)) return RCABEAM_ERROR_ARGUMENT;
  EnsembleArgs p = make_args(rc, cr, nullptr, xe, ye, angles, starts, nullptr, nullptr, nullptr, nullptr, ns, nc, na,
                             nt, nv, c, fs, fd, fn);
  const hipError_t e = launch_opw_adjoint_grid(p, x, z, y, static_cast<int>(nx), static_cast<int>(nz),
                                               static_cast<int>(ny), static_cast<const float2*>(grad_iq),
                                               static_cast<float2*>(grad_rc), static_cast<float2*>(grad_cr), grad_c,
                                               workspace);
  if (e == hipErrorNotSupported) return RCABEAM_ERROR_ARGUMENT;
  return to_status(e);
}
