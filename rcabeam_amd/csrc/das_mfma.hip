// OPW delay-and-sum on MFMA for 16-bit channel storage on a Cartesian grid:
// the 16-bit path of rcabeam_opw_grid_device (reference
// src/rcabeam/ensemble.cu:341-454 computes the same sums with per-tap
// float32 gathers).
//
// For one (aperture, channel, angle) a voxel adds P0 S[i0, t] + P1 S[i0+1, t]
// to its IQ, P0 = (1 - w) e^{i phi}, P1 = w e^{i phi}.  For a group of 16
// voxels (a 4 x 4 lateral patch at one depth) the rows i0, i0 + 1 of the
// group lie in a short window (4.1 rows on average for 128+128 elements, 11
// angles over 4-20 mm), so the update is a small dense product
//   OUT (16 voxels x 32 frames) += C (16 voxels x 8 rows) S (8 rows x 32 frames)
// with two non-zero coefficients per voxel.  v_mfma_f32_32x32x16_{f16,bf16}
// computes it in real form: M = 32 = (voxel, Re / Im of the output),
// N = 32 frames, K = 16 = (window row, Re / Im of the sample); lanes 0-31
// hold window rows 0-3 and lanes 32-63 rows 4-7.  Products of 16-bit values
// are exact in the float32 accumulator:
//   kCoefSingle : one 16-bit coefficient per sample (1 MFMA per K-step and block)
//   kCoefSplit  : hi + lo FP16 coefficients, float32-accurate (2 MFMAs)
// SPLIT channels (complex64 kept as scaled FP16 hi + lo word pairs, 8 bytes
// per sample, see split_exponent) run the float32-accurate form of complex64
// channels: C_hi S_hi + C_lo S_hi + C_hi S_lo (3 MFMAs; C_lo S_lo is below
// 2^-22 of the product), and the epilogue undoes the power-of-two scale.
//
// Work decomposition: one 256-thread workgroup (4 wave64s, two per CU) owns
// an 8 x 8 lateral patch at one depth and all frames (tiles in XCD-aware
// 4 x 4 x 4 blocks); wave g owns the 16 voxels of group g and every
// 32-frame block (NB accumulator blocks of 16 floats per lane).
//
// On a grid the delay separates: the receive term of RC depends on (x, z,
// channel), the transmit term on (y, z, angle), and the other way round for
// CR.  A tile therefore needs 8 receive terms per channel and 8 transmit terms
// per angle (float64, split into integer sample, float32 fraction and unit
// phasor).  Transmit terms are formed once per tile into LDS; each wave forms
// the 8 receive terms of a channel itself.  Because the first tap row is
// monotonic in both terms, the row window of a set of voxels is exactly
// [first row (min receive + min transmit), last row (max receive + max
// transmit)] clipped to the closed sample window, so the windows of the tile
// (staging) and of each group (MFMA K-steps) come from min / max over 8 or 4
// terms, without a tap table.  Each lane combines its A-operand voxel's two
// terms into its tap.  Channels whose dynamic aperture misses the tile are
// skipped.
//
// Windows of consecutive angles are packed back to back into stages; a
// window's rows are one contiguous block in the packed (channels, angles,
// samples, frames) layout and in LDS (row stride nt), so a stage is copied
// with 1 KiB global_load_lds_dwordx4 instructions regardless of row
// boundaries, double-buffered across angles and channels, one barrier per
// stage.  Each window is copied by one wave in a loop of whole chunks (no
// lane mask) and one masked tail chunk, and the next angle's transmit term
// is read from LDS while the current angle is consumed.  B operands of frames past nt read the next row or zero padding and
// only feed accumulator columns that are never stored.
#include "grid_tile.hpp"
#include "rcabeam_mi355x.h"

#include <algorithm>

namespace rcb {
namespace {

constexpr int kMThreads = 256;
constexpr int kMWaves = kMThreads / kWave;  // one wave per 16-voxel group
constexpr int kTileSlots = 64;
constexpr int kBlockFrames = 32;  // N of one MFMA
constexpr int kMaxBlocks = 7;     // nt <= 224
constexpr int kKRows = 8;         // window rows of one K-step
constexpr int kMaxRowsM = 64;
constexpr int kMinRowsM = 12;
constexpr int kLdsBytesM = 80 * 1024;  // two workgroups per CU

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ f32x16 mfma(u32x4 a, u32x4 b, f32x16 c);
template <>
__device__ __forceinline__ f32x16 mfma<__half2>(u32x4 a, u32x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0,
                                                0);
}
template <>
__device__ __forceinline__ f32x16 mfma<bf16x2>(u32x4 a, u32x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}

// Coefficient word of an imaginary output row from that of the real row:
// (P.re, -P.im) -> (P.im, P.re), i.e. swap the halves, negate the new low half.
__device__ __forceinline__ uint32_t imag_word(uint32_t w) {
  return __builtin_amdgcn_alignbit(w, w, 16) ^ 0x00008000u;
}

// A operand of one lane: coefficient words of its 4 window rows, c0 at row
// rel (the tap's first row relative to the lane's first row) and c1 at
// rel + 1; rel outside [-1, 3] gives zeros.
__device__ __forceinline__ u32x4 place(int rel, uint32_t c0, uint32_t c1) {
  u32x4 a;
  a[0] = rel == 0 ? c0 : (rel == -1 ? c1 : 0u);
  a[1] = rel == 1 ? c0 : (rel == 0 ? c1 : 0u);
  a[2] = rel == 2 ? c0 : (rel == 1 ? c1 : 0u);
  a[3] = rel == 3 ? c0 : (rel == 2 ? c1 : 0u);
  return a;
}

struct MLayout {  // byte offsets inside the dynamic LDS allocation
  int psum, win, buffer_words, bytes;
};

// Tile tables (grid_tile.hpp) at 0; psum: float[64]; win: two buffers of
// max_rows rows x nt samples (wps words each), each followed by zero padding
// that K-steps at the end of a stage and frame blocks past nt may read.
__host__ __device__ inline MLayout mlayout(int nc, int nt, int nb, int max_rows, int wps) {
  MLayout l;
  l.psum = tile_table_bytes(nc);
  l.win = (l.psum + kTileSlots * 4 + 15) & ~15;
  l.buffer_words = (((max_rows + kKRows) * nt + kBlockFrames * nb) * wps + 3) & ~3;
  l.bytes = l.win + 2 * l.buffer_words * 4;
  return l;
}

template <typename T, int CP, int NB, bool SPLIT>
__global__ __launch_bounds__(kMThreads) __attribute__((amdgpu_waves_per_eu(2))) void das_opw_mfma(GridArgs p) {
  constexpr int CW = CP == kCoefSplit ? 4 : 2;
  constexpr int WPS = SPLIT ? 2 : 1;  // 32-bit words per staged sample
  using W = uint32_t;
  extern __shared__ __align__(16) unsigned char lds[];
  const MLayout layout = mlayout(p.nc, p.nt, NB, p.max_rows, WPS);
  float* psum = reinterpret_cast<float*>(lds + layout.psum);  // [64]
  W* win = reinterpret_cast<W*>(lds + layout.win);            // [2][buffer_words]
  const int buffer = layout.buffer_words;

  GridTile<kMThreads> tile(p, lds);
  if (!tile.valid()) return;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int g = uniform(tid / kWave);  // voxel group (tile quarter) of this wave, wave-uniform
  const int col = lane & 31;      // frame column of the B operand and the output
  const int kg = lane >> 5;       // window rows 4 kg .. 4 kg + 3 of a K-step
  const bool imag_row = (col & 1) != 0;
  const int nt = p.nt;
  const int nq = 2 * p.nc;
  const int max_rows = p.max_rows;
  // A-operand voxel of this lane: group g's (vx, vy) = ((col / 2) / 4, (col / 2) % 4).
  const int ax = 4 * (g >> 1) + (col >> 3);
  const int ay = 4 * (g & 1) + ((col >> 1) & 3);

  {  // both window buffers and their padding zeroed once
    uint4* w4 = reinterpret_cast<uint4*>(win);
    const int n = 2 * buffer / 4;
    for (int i = tid; i < n; i += kMThreads) w4[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  tile.build_tables();  // ends with a barrier
  auto setup = [&](int q) { return tile.setup(q, ax, ay, g); };
  auto next_live = [&](int q) { return tile.next_live(q); };
  auto plane_of = [&](const Channel& ch) {
    const W* data = static_cast<const W*>(ch.cfg == 0 ? p.rc : p.cr);
    return data + static_cast<size_t>(ch.q - ch.cfg * p.nc) * p.na * p.ns * nt * WPS;
  };
  const size_t plane_size = static_cast<size_t>(p.ns) * nt;

  // Copy `words` 32-bit words (a multiple of 4) from HBM to LDS, 1 KiB per
  // instruction: chunks first, first + step, ... of this wave; whole chunks
  // need no lane mask, only the last (partial) chunk does.
  auto copy_words = [&](const W* src, W* dst, int words, int first, int step) {
    if (p.vec) {
      const int full = words >> 8;  // whole 1 KiB chunks
      const W* s = src + lane * 4;
      for (int i = first; i < full; i += step)
        __builtin_amdgcn_global_load_lds(
            reinterpret_cast<const void*>(s + i * 256),
            reinterpret_cast<__attribute__((address_space(3))) void*>(reinterpret_cast<uintptr_t>(dst + i * 256)),
            16, 0, 0);
      if ((words & 255) && (full - first) % step == 0 && lane * 4 < (words & 255))
        __builtin_amdgcn_global_load_lds(
            reinterpret_cast<const void*>(s + full * 256),
            reinterpret_cast<__attribute__((address_space(3))) void*>(reinterpret_cast<uintptr_t>(dst + full * 256)),
            16, 0, 0);
    } else {
      for (int i = first * kWave + lane; i < words; i += step * kWave) dst[i] = src[i];
    }
  };
  // Copy a stage into buf: windows of angles [a0, a1) back to back (or one
  // chunk of a long window), contiguous in HBM and in LDS (row stride nt),
  // 1 KiB global_load_lds_dwordx4 instructions.  Each window is copied whole
  // by one wave (angle a0 + g, a0 + g + 4, ... by wave g); its LDS offset is
  // the exclusive prefix of the row counts over lanes a0 .. a - 1 (lane a
  // holds angle a's window; na <= 16 keeps the scan inside one DPP row).  A
  // chunk of a long window is spread over all waves.
  auto issue = [&](const Channel& ch, const Stage& s, W* buf) {
    const W* plane0 = plane_of(ch);
    const int a0 = uniform(s.a0), a1 = uniform(s.a1), sr0 = uniform(s.r0), srows = uniform(s.rows);
    if (sr0 >= 0) {
      copy_words(plane0 + (a0 * plane_size + static_cast<size_t>(sr0) * nt) * WPS, buf, srows * nt * WPS, g,
                 kMWaves);
      return;
    }
    const int rows_l = (lane >= a0 && lane < a1 && ch.w.lo <= ch.w.hi) ? ch.w.hi - ch.w.lo + 1 : 0;
    int incl = rows_l;
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xf, 0xf, true);  // row_shr:1
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xf, 0xf, true);  // row_shr:2
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xf, 0xf, true);  // row_shr:4
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xf, 0xf, true);  // row_shr:8
    const int excl = incl - rows_l;
    for (int a = a0 + g; a < a1; a += kMWaves) {
      const int2 b = ch.w.at(a);
      if (b.x > b.y) continue;
      const int off = __builtin_amdgcn_readlane(excl, a);
      copy_words(plane0 + (a * plane_size + static_cast<size_t>(b.x) * nt) * WPS, buf + off * nt * WPS,
                 (b.y - b.x + 1) * nt * WPS, 0, 1);
    }
  };

  f32x16 acc[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[j][e] = 0.f;

  // This wave's group, one stage: K-steps of 8 window rows per angle.
  auto consume = [&](const Channel& ch, const Stage& cur, const W* base) {
    // Stage fields are wave-uniform; saying so keeps the angle and K-step loops scalar.
    const int a0 = uniform(cur.a0), a1 = uniform(cur.a1), cr0 = uniform(cur.r0), crows = uniform(cur.rows);
    const int t_lo = cr0 >= 0 ? cr0 : -kBig;
    const int t_hi = cr0 >= 0 && !cur.last ? cr0 + max_rows - 2 : kBig;
    const int tpos = ch.cfg == 0 ? ay : ax;  // transmit position of the A-operand voxel
    int off = 0;
    Term tx_next = tile.transmit(ch.cfg, a0, tpos);
    for (int a = a0; a < a1; ++a) {
      // The next angle's transmit term is read while this angle is consumed.
      const Term tx = tx_next;
      if (a + 1 < a1) tx_next = tile.transmit(ch.cfg, a + 1, tpos);
      int first, row_off, lim_lo, lim_hi;
      if (cr0 >= 0) {  // one chunk of a long window
        first = cr0;
        row_off = 0;
        lim_lo = cr0;
        lim_hi = cr0 + crows - 1;
      } else {  // whole windows packed back to back
        const int2 bw = ch.w.at(a);
        if (bw.x > bw.y) continue;
        first = bw.x;
        row_off = off;
        off += bw.y - bw.x + 1;
        lim_lo = bw.x;
        lim_hi = bw.y;
      }
      const int2 gw = ch.g.at(a);
      const int lo = max(gw.x, lim_lo);
      const int hi = min(gw.y, lim_hi);
      if (lo > hi) continue;
      const Tap tp = tap_of(ch.rx, tx, p.ns);
      const bool take = tp.i0 >= 0 && tp.i0 >= t_lo && tp.i0 <= t_hi;
      const float u = 1.0f - tp.w;
      const float p0r = u * tp.c, p0i = u * tp.s, p1r = tp.w * tp.c, p1i = tp.w * tp.s;
      uint32_t c0 = pack2<T>(p0r, -p0i), c1 = pack2<T>(p1r, -p1i), l0 = 0, l1 = 0;
      if constexpr (CW == 4) {
        const float2 h0 = widen<T>(c0), h1 = widen<T>(c1);
        l0 = pack2<T>(p0r - h0.x, -p0i - h0.y);
        l1 = pack2<T>(p1r - h1.x, -p1i - h1.y);
        l0 = imag_row ? imag_word(l0) : l0;
        l1 = imag_row ? imag_word(l1) : l1;
      }
      c0 = imag_row ? imag_word(c0) : c0;
      c1 = imag_row ? imag_word(c1) : c1;
      for (int r = lo; r <= hi; r += kKRows) {
        const int rel = take ? tp.i0 - r - 4 * kg : 64;
        const u32x4 ahi = place(rel, c0, c1);
        // 0 <= row < 2^24: a 24-bit multiply gives the sample offset.
        const uint32_t s0 = __umul24(static_cast<uint32_t>(r - first + row_off + 4 * kg), static_cast<uint32_t>(nt)) + col;
        if constexpr (SPLIT) {
          // One 8-byte read per (row, frame): the hi and lo words of the sample.  Block by block, so
          // only one block's operands are live (the three MFMAs of a block cover the next block's reads).
          const uint2* b0 = reinterpret_cast<const uint2*>(base) + s0;
          const u32x4 alo = place(rel, l0, l1);
          uint2 w[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) w[k] = b0[k * nt];
#pragma unroll
          for (int j = 0; j < NB; ++j) {
            const u32x4 bh = {w[0].x, w[1].x, w[2].x, w[3].x};
            const u32x4 bl = {w[0].y, w[1].y, w[2].y, w[3].y};
            if (j + 1 < NB) {
#pragma unroll
              for (int k = 0; k < 4; ++k) w[k] = b0[k * nt + kBlockFrames * (j + 1)];
            }
            acc[j] = mfma<T>(ahi, bh, acc[j]);
            acc[j] = mfma<T>(alo, bh, acc[j]);
            acc[j] = mfma<T>(ahi, bl, acc[j]);
            // One block's reads in flight behind the current block's MFMAs; hoisting every block's reads
            // would spill registers.
            __builtin_amdgcn_sched_barrier(0);
          }
        } else {
          const W* b0 = base + s0;
          u32x4 bv[NB];
#pragma unroll
          for (int j = 0; j < NB; ++j) {
            bv[j][0] = b0[kBlockFrames * j];
            bv[j][1] = b0[nt + kBlockFrames * j];
            bv[j][2] = b0[2 * nt + kBlockFrames * j];
            bv[j][3] = b0[3 * nt + kBlockFrames * j];
          }
#pragma unroll
          for (int j = 0; j < NB; ++j) acc[j] = mfma<T>(ahi, bv[j], acc[j]);
          if constexpr (CW == 4) {
            const u32x4 alo = place(rel, l0, l1);
#pragma unroll
            for (int j = 0; j < NB; ++j) acc[j] = mfma<T>(alo, bv[j], acc[j]);
          }
        }
      }
    }
  };

  // ---- stages of all live channels, one barrier per stage
  Stage cur{0, 0, -1, 0, true, false};
  Channel ch;
  ch.q = -1;
  for (int q = next_live(-1); q < nq; q = next_live(q)) {
    ch = setup(q);
    cur = plan_stage_lanes(ch.w, 0, p.na, max_rows);
    if (cur.valid) break;
  }
  int b = 0;
  if (cur.valid) issue(ch, cur, win);
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  __syncthreads();
  while (cur.valid) {
    Stage nxt = next_stage_lanes(cur, ch.w, p.na, max_rows);
    Channel nch = ch;
    bool crossing = false;
    if (!nxt.valid) {
      for (int q = next_live(ch.q); q < nq; q = next_live(q)) {
        nch = setup(q);
        nxt = plan_stage_lanes(nch.w, 0, p.na, max_rows);
        if (nxt.valid) break;
      }
      crossing = nxt.valid;
    }
    if (nxt.valid) issue(crossing ? nch : ch, nxt, win + (b ^ 1) * buffer);
    consume(ch, cur, win + b * buffer);
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __syncthreads();
    b ^= 1;
    if (crossing) ch = nch;
    cur = nxt;
  }

  // ---- epilogue.  acc[j][e] is output row 8 (e / 4) + 4 kg + e % 4 = (voxel
  // 4 (e / 4) + 2 kg + (e / 2) % 2 of the group, Re / Im = e % 2) at frame
  // 32 j + col; group voxel u sits at (4 (g / 2) + u / 4, 4 (g % 2) + u % 4).
  auto voxel_id = [&](int u) { return tile.voxel(4 * (g >> 1) + (u >> 2), 4 * (g & 1) + (u & 3)); };
  long long vid[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) vid[k] = voxel_id(4 * (k >> 1) + 2 * kg + (k & 1));
  float pw[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) pw[k] = 0.f;
  // Split channels: undo the power-of-two channel scale (exact).
  const int unscale = SPLIT ? -split_exponent(p.absmax) : 0;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int t = kBlockFrames * j + col;
    if (t >= nt) continue;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      float2 s = make_float2(acc[j][2 * k], acc[j][2 * k + 1]);
      if constexpr (SPLIT) s = make_float2(ldexpf(s.x, unscale), ldexpf(s.y, unscale));
      pw[k] = fmaf(s.x, s.x, fmaf(s.y, s.y, pw[k]));
      if (p.signal && vid[k] >= 0) {
        const long long index = vid[k] * nt + t;
        if (p.out_storage == RCABEAM_STORAGE_FP16)
          store_sample(reinterpret_cast<__half2*>(p.signal) + index, s);
        else if (p.out_storage == RCABEAM_STORAGE_BF16)
          store_sample(reinterpret_cast<bf16x2*>(p.signal) + index, s);
        else
          store_sample(reinterpret_cast<float2*>(p.signal) + index, s);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) pw[k] += __shfl_xor(pw[k], off, kWave);
  if (col == 0) {
#pragma unroll
    for (int k = 0; k < 8; ++k) psum[16 * g + 4 * (k >> 1) + 2 * kg + (k & 1)] = pw[k];
  }
  __syncthreads();
  if (tid < kTileSlots) {
    const int gg = tid >> 4, u = tid & 15;
    const long long v = tile.voxel(4 * (gg >> 1) + (u >> 2), 4 * (gg & 1) + (u & 3));
    if (v >= 0) {
      p.pd[2 * v] = psum[tid] / static_cast<float>(nt);
      p.pd[2 * v + 1] = 0.f;
    }
  }
}

template <typename T, int CP, int NB, bool SPLIT>
hipError_t launch_shape(GridArgs a) {
  constexpr int WPS = SPLIT ? 2 : 1;
  const int fixed = mlayout(a.nc, a.nt, NB, 0, WPS).bytes;
  a.max_rows = std::min(kMaxRowsM, (kLdsBytesM - fixed) / (2 * a.nt * 4 * WPS));
  if (a.max_rows < kMinRowsM) return hipErrorNotSupported;
  const int bytes = mlayout(a.nc, a.nt, NB, a.max_rows, WPS).bytes;
  if (bytes > kLdsBytesM) return hipErrorNotSupported;
  const long long grid = (a.tiles + 7) / 8 * 8;
  if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
  // 1 KiB copies need 16-byte aligned planes and rows of a multiple of 16 bytes.
  a.vec = reinterpret_cast<uintptr_t>(a.rc) % 16 == 0 && reinterpret_cast<uintptr_t>(a.cr) % 16 == 0 &&
          (a.nt * WPS) % 4 == 0;
  auto kernel = das_opw_mfma<T, CP, NB, SPLIT>;
  const hipError_t e = allow_dynamic_lds(reinterpret_cast<const void*>(kernel), kLdsBytesM);
  if (e != hipSuccess) return e;
  kernel<<<dim3(static_cast<unsigned>(grid)), kMThreads, bytes>>>(a);
  return hipGetLastError();
}

template <typename T, int CP, bool SPLIT = false>
hipError_t launch_blocks(const GridArgs& a) {
  const int blocks = (a.nt + kBlockFrames - 1) / kBlockFrames;
  if (blocks <= 1) return launch_shape<T, CP, 1, SPLIT>(a);
  if (blocks <= 2) return launch_shape<T, CP, 2, SPLIT>(a);
  if (blocks <= 4) return launch_shape<T, CP, 4, SPLIT>(a);
  return launch_shape<T, CP, kMaxBlocks, SPLIT>(a);
}

}  // namespace

hipError_t launch_opw_grid_mfma(const EnsembleArgs& p, const float* x, const float* z, const float* y, int nx, int nz,
                                int ny, int channel_storage, int coefficients, int signal_storage,
                                const unsigned* absmax) {
  if (p.na > kGridMaxAngles || p.nt > kBlockFrames * kMaxBlocks) return hipErrorNotSupported;
  if (static_cast<long long>(p.ns) * p.nt > 0x7fffffffLL) return hipErrorNotSupported;
  if (signal_storage < RCABEAM_STORAGE_C64 || signal_storage > RCABEAM_STORAGE_BF16) return hipErrorInvalidValue;
  GridArgs a = make_grid_args(p, x, z, y, nx, nz, ny, signal_storage);
  a.absmax = absmax;
  switch (channel_storage) {
    case RCABEAM_STORAGE_FP16_SPLIT:
      if (!absmax) return hipErrorInvalidValue;
      return launch_blocks<__half2, kCoefSplit, true>(a);
    case RCABEAM_STORAGE_FP16:
      if (coefficients == kCoefSingle) return launch_blocks<__half2, kCoefSingle>(a);
      if (coefficients == kCoefSplit) return launch_blocks<__half2, kCoefSplit>(a);
      return hipErrorNotSupported;
    case RCABEAM_STORAGE_BF16:
      if (coefficients == kCoefSingle) return launch_blocks<bf16x2, kCoefSingle>(a);
      return hipErrorNotSupported;
    default:
      return hipErrorNotSupported;
  }
}

}  // namespace rcb
