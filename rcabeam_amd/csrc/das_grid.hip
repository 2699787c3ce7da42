// OPW delay-and-sum of complex64 channels on a Cartesian grid, float32 VALU
// arithmetic: the complex64 OPW path of rcabeam_opw_grid_device (reference
// src/rcabeam/ensemble.cu:341-454 computes the same sums with per-tap
// float32 gathers).
//
// Work decomposition (one 512-thread workgroup = 8 wave64s, two per CU): a
// workgroup owns an 8 x 8 lateral patch of voxels at one depth and all
// frames (tiles in XCD-aware 4 x 4 x 4 blocks).  Wave w owns a 2 (x) x 4 (y)
// block of voxels, so the aperture mask of an RC channel (a function of x)
// and of a CR channel (of y) each vary over only 2 or 4 of a wave's voxels:
// fewer idle lanes than 8 positions along y, and less idle waves at the stage
// barriers than one x row per wave (1 % faster than either).  The 8 lanes
// l = 8 g + f of the wave's voxel g own the frame pairs
// (2 f + 16 j, 2 f + 16 j + 1), j < FPP, so one ds_read_b128 of a wave reads
// 128 contiguous bytes of one sample row per voxel.  With TAIL the last block
// holds at most 8 frames and lane f owns the single frame 16 (FPP - 1) + f
// (ds_read_b64), so 200 frames cost 200 frame updates per lane group, not
// 208.  Tile tables, analytic windows, live-channel skipping and stage
// planning are those of grid_tile.hpp; windows of consecutive
// angles are packed into stages copied from the packed (channels, angles,
// samples, frames) layout into LDS rows of nt samples with 1 KiB
// global_load_lds_dwordx4 instructions, double-buffered across angles and
// channels, one barrier per stage.
//
// Stages are planned once per workgroup, not once per wave: wave k % 8 plans
// stage k + 2 (next channel, receive terms, windows, packing, LDS offsets)
// into a ring of three LDS descriptors while stage k is consumed and stage
// k + 1 is copied; the other waves read the descriptors.  Repeating the plan
// in all 8 waves cost ~18 % of the kernel's VALU instructions.
//
// Per tap and frame: v = a + w (b - a), acc += v e^{i phi} (8 FMA), with the
// tap (row, weight, phasor) of the lane's own voxel combined from its
// receive and transmit terms exactly as combine() does.  The tap is formed
// only as far as the branch needs it: an inactive term carries kInactive
// (-2^30, active ones stay within +/- 2^28), so any inactive term already
// makes the row negative, and the phasor is formed only for valid taps.
//
// Single-angle launches (na == 1, MULTI): a channel has one window, so a
// stage packs the windows of up to 4 consecutive live channels of one
// aperture instead (a window longer than a stage is still copied in chunks).
// The receive terms of a stage's channels (8 tile positions each) go to a
// small LDS slot when the stage is planned, so reading the stage needs
// neither registers held across it nor a second receive-term fill.
//
// APERTURES (method 5): RC channels come first, so when the first CR channel
// is reached the sums are the RC IQ; they are stored to plane 0 of the
// (voxel, aperture, frame) output and restarted, and the CR IQ goes to
// plane 1 at the end.
#include "grid_tile.hpp"
#include "rcabeam_mi355x.h"

#include <algorithm>

namespace rcb {
namespace {

constexpr int kVThreads = 512;
constexpr int kVWaves = kVThreads / kWave;
constexpr int kVTileSlots = 64;
constexpr int kVMaxRows = 48;
constexpr int kVMinRows = 8;
constexpr int kVLdsBytes = 80 * 1024;  // two workgroups per CU
constexpr int kMaxPairs = 13;          // nt <= 208
// Each stage buffer is its own LDS variable (36 KiB).  The compiler can then
// tell that LDS reads of one buffer do not touch the copies in flight into
// the other, and does not wait for those copies before the reads.
constexpr int kVStageWords = 9216;

constexpr int kStageChannels = 4;  // channel windows per stage of a single-angle launch

// A planned stage, shared by the waves of a workgroup through LDS: one wave
// plans each stage (two ahead of the one being consumed) and every wave reads
// the plan instead of repeating it.
struct StageDesc {
  int q;        // channel (-1: no further stage)
  int a0, a1;   // angles [a0, a1)
  int r0;       // >= 0: one chunk of a long window, rows [r0, r0 + rows) of angle a0
  int rows;
  int last;     // chunk stage: the last chunk of its window
  int pad0, pad1;
  int lo[16];   // tile window of each angle of channel q (lo > hi: empty)
  int hi[16];
  int off[16];  // LDS row of each window's first row (whole-window stages)
  int base[16]; // LDS row = sample row - base (whole-window stages)
  Term rx[8];   // receive terms of channel q at tile positions 0 .. 7
};
constexpr int kDescBytes = static_cast<int>(sizeof(StageDesc));
constexpr int kDescs = 3;  // consumed, being copied, being planned

// Dynamic LDS: tile tables at 0, then psum: float[64], then (single-angle
// launches) two stage slots: receive terms Term[kStageChannels][8] and row
// shifts int[kStageChannels] each.
constexpr int kSlotTermBytes = kStageChannels * kGridPatch * 16;
constexpr int kSlotBytes = kSlotTermBytes + kStageChannels * 4;
__host__ __device__ inline int psum_offset(int nc) { return tile_table_bytes(nc); }
__host__ __device__ inline int slot_offset(int nc) { return psum_offset(nc) + kVTileSlots * 4; }
__host__ __device__ inline int desc_offset(int nc) { return (slot_offset(nc) + 15) & ~15; }
__host__ __device__ inline int dynamic_bytes(int nc, bool multi) {
  return multi ? slot_offset(nc) + 2 * kSlotBytes : desc_offset(nc) + kDescs * kDescBytes;
}

// Rows of nt complex64 samples a stage buffer holds, with room for the frames
// the last row's reads cover past nt: pair blocks reach frame 16 fpp, or
// 16 (fpp - 1) + 8 with a tail block (none past nt = 200).
__host__ __device__ inline int stage_rows(int nt, int fpp, bool tail) {
  const int span = tail ? 16 * (fpp - 1) + 8 : 16 * fpp;
  return (kVStageWords / 2 - (span > nt ? span - nt : 0)) / nt;
}

// A stage of a single-angle launch: n channel windows (or one chunk of a
// long window, taps filtered to rows [t_lo, t_hi]) of aperture cfg.
struct MultiStage {
  int n, cfg, t_lo, t_hi;
  bool valid;
};

template <int FPP, bool APERTURES, bool MULTI, bool TAIL>
__global__ __launch_bounds__(kVThreads) __attribute__((amdgpu_waves_per_eu(4))) void das_opw_grid_valu(GridArgs p) {
  constexpr int kPairs = TAIL ? FPP - 1 : FPP;      // full pair blocks
  constexpr int kAcc = 2 * kPairs + (TAIL ? 1 : 0);  // frames per lane
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ __align__(16) float2 stage_a[kVStageWords / 2];
  __shared__ __align__(16) float2 stage_b[kVStageWords / 2];
  float* psum = reinterpret_cast<float*>(lds + psum_offset(p.nc));  // [64]

  GridTile<kVThreads> tile(p, lds);
  if (!tile.valid()) return;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = uniform(tid / kWave);  // wave-uniform: loops over it stay scalar
  // This lane's voxel (tile-local x, y): the wave's 2 x 4 block, voxel g = lane / 8 at (g % 2, g / 2) in it.
  const int vx = 2 * (wave & 3) + ((lane >> 3) & 1);
  const int vy = 4 * (wave >> 2) + ((lane >> 3) >> 1);
  const int fl = lane & 7;       // frame pairs 2 fl + 16 j
  const int nt = p.nt;
  const int nq = 2 * p.nc;
  const int max_rows = p.max_rows;

  // Both stage buffers zeroed once (pairs past nt then read zeros or finite samples).
  for (int i = tid; i < kVStageWords / 4; i += kVThreads) {
    reinterpret_cast<uint4*>(stage_a)[i] = make_uint4(0u, 0u, 0u, 0u);
    reinterpret_cast<uint4*>(stage_b)[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  tile.build_tables();  // ends with a barrier

  const size_t plane_size = static_cast<size_t>(p.ns) * nt;
  // Receive positions of this lane's voxel: x for RC, y for CR.
  auto setup = [&](int q) __attribute__((always_inline)) { return tile.setup(q, vx, vy, 0); };

  // Rows [0, rows) of src (nt samples each) to dst, 1 KiB per wave
  // instruction spread over the waves, asynchronous.
  auto copy_rows = [&](const float2* src, float2* dst, int rows) __attribute__((always_inline)) {
    const int samples = rows * nt;
    const int chunks = (samples + 127) >> 7;  // 128 samples = 1 KiB
    for (int i = wave; i < chunks; i += kVWaves) {
      const int so = i * 128 + lane * 2;
      if (so < samples)
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(src + so),
                                         (__attribute__((address_space(3))) void*)(dst + i * 128), 16, 0, 0);
    }
  };
  // Rows [0, rows) of src to dst, all issued by this wave.
  auto copy_rows_wave = [&](const float2* src, float2* dst, int rows) __attribute__((always_inline)) {
    const int samples = rows * nt;
    const int chunks = (samples + 127) >> 7;
    for (int i = 0; i < chunks; ++i) {
      const int so = i * 128 + lane * 2;
      if (so < samples)
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(src + so),
                                         (__attribute__((address_space(3))) void*)(dst + i * 128), 16, 0, 0);
    }
  };
  float2 acc[kAcc];
#pragma unroll
  for (int j = 0; j < kAcc; ++j) acc[j] = make_float2(0.f, 0.f);
  // Frame of accumulator k of this lane.
  auto frame_of = [&](int k) __attribute__((always_inline)) {
    return (TAIL && k == 2 * kPairs) ? 16 * kPairs + fl : 2 * fl + 16 * (k >> 1) + (k & 1);
  };

  const uint32_t row_bytes = static_cast<uint32_t>(nt) * 8u;
  // One tap over all frames from staged row r0 (0 <= r0 < 2^24 here, so a
  // 24-bit multiply gives its byte offset); row r0 + 1 only where the weight
  // is non-zero.
  auto accumulate = [&](const Tap& tp, uint32_t r0, const float2* base) __attribute__((always_inline)) {
    const unsigned char* row0 = reinterpret_cast<const unsigned char*>(base) + __umul24(r0, row_bytes);
    const unsigned next = tp.w != 0.f ? row_bytes : 0u;
    const unsigned char* row = row0 + 16 * fl;
    const float4* s0 = reinterpret_cast<const float4*>(row);
    const float4* s1 = reinterpret_cast<const float4*>(row + next);
    if constexpr (TAIL) {
      const float2* t0 = reinterpret_cast<const float2*>(row0 + 128 * kPairs + 8 * fl);
      const float2* t1 = reinterpret_cast<const float2*>(row0 + 128 * kPairs + 8 * fl + next);
      const float2 va = *t0, vb = *t1;  // frame 16 kPairs + fl
      const float ax = fmaf(tp.w, vb.x - va.x, va.x), ay = fmaf(tp.w, vb.y - va.y, va.y);
      acc[2 * kPairs].x = fmaf(tp.c, ax, fmaf(-tp.s, ay, acc[2 * kPairs].x));
      acc[2 * kPairs].y = fmaf(tp.c, ay, fmaf(tp.s, ax, acc[2 * kPairs].y));
    }
    // Software-pipelined: the rows of pair block j + 1 are read while block j
    // is computed (a wave keeps two blocks of LDS reads in flight; the
    // scheduling barriers stop the compiler from sinking the reads back).
    float4 va = s0[0], vb = s1[0];  // frames 2 fl, 2 fl + 1
#pragma unroll
    for (int j = 0; j < kPairs; ++j) {
      float4 na, nb;
      if (j + 1 < kPairs) {
        na = s0[8 * (j + 1)];
        nb = s1[8 * (j + 1)];
      }
      __builtin_amdgcn_sched_barrier(0);
      const float ax = fmaf(tp.w, vb.x - va.x, va.x), ay = fmaf(tp.w, vb.y - va.y, va.y);
      const float bx = fmaf(tp.w, vb.z - va.z, va.z), by = fmaf(tp.w, vb.w - va.w, va.w);
      acc[2 * j].x = fmaf(tp.c, ax, fmaf(-tp.s, ay, acc[2 * j].x));
      acc[2 * j].y = fmaf(tp.c, ay, fmaf(tp.s, ax, acc[2 * j].y));
      acc[2 * j + 1].x = fmaf(tp.c, bx, fmaf(-tp.s, by, acc[2 * j + 1].x));
      acc[2 * j + 1].y = fmaf(tp.c, by, fmaf(tp.s, bx, acc[2 * j + 1].y));
      __builtin_amdgcn_sched_barrier(0);
      if (j + 1 < kPairs) {
        va = na;
        vb = nb;
      }
    }
  };

  // Apertures: store the sums as plane `cfg` of this lane's voxel and restart them.
  const long long v = tile.voxel(vx, vy);
  auto store_plane = [&](int cfg) __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < kAcc; ++k) {
      const int t = frame_of(k);
      if (t < nt && v >= 0) reinterpret_cast<float2*>(p.signal)[(2 * v + cfg) * nt + t] = acc[k];
      acc[k] = make_float2(0.f, 0.f);
    }
  };
  bool rc_stored = false;

  if constexpr (MULTI) {
    // ---- single angle: stages of up to kStageChannels channel windows.
    Term* slot_rx = reinterpret_cast<Term*>(lds + slot_offset(p.nc));  // [2][kStageChannels][8]
    int* slot_shift = reinterpret_cast<int*>(lds + slot_offset(p.nc) + 2 * kSlotTermBytes);  // [2][kStageChannels]
    // Next piece to stage: channel cq from row cr (cr < 0: its whole window).
    int cq = tile.next_live(-1), cr = -1;
    // Plan the stage at (cq, cr), issue its copies into buf and record its
    // receive terms and row shifts in slot s (wave 0 writes; every wave plans
    // the same stage from tile-wide windows).  Advances (cq, cr).
    auto plan_issue = [&](float2* buf, int s) __attribute__((always_inline)) {
      MultiStage st{0, 0, -kBig, kBig, false};
      int total = 0;
      auto record = [&](int k, int q, int lo) __attribute__((always_inline)) {
        const int first = uniform(8 * (q - tile.rx_base));
        if (wave == 0) {
          if (lane >= first && lane < first + 8) slot_rx[(s * kStageChannels + k) * 8 + (lane - first)] = tile.rx_term;
          if (lane == 0) slot_shift[s * kStageChannels + k] = total - lo;
        }
      };
      const float2* const planes[2] = {static_cast<const float2*>(p.rc), static_cast<const float2*>(p.cr)};
      auto plane = [&](int q) __attribute__((always_inline)) {
        const int cfg = q >= p.nc ? 1 : 0;
        return (cfg == 0 ? planes[0] : planes[1]) + static_cast<size_t>(q - cfg * p.nc) * plane_size;
      };
      if (cr >= 0) {  // next chunk of a window longer than a stage
        const Channel ch = setup(cq);
        const int hi = ch.w.at(0).y;
        const bool last = cr + max_rows - 1 >= hi;
        const int rows = last ? hi + 1 - cr : max_rows;
        record(0, cq, cr);
        copy_rows(plane(cq) + static_cast<size_t>(cr) * nt, buf, rows);
        st = MultiStage{1, ch.cfg, cr, last ? kBig : cr + max_rows - 2, true};
        if (last) {
          cq = tile.next_live(cq);
          cr = -1;
        } else {
          cr += max_rows - 1;
        }
        return st;
      }
      // Receive terms of the stage's channels filled before its first copy:
      // a fill's coordinate loads would otherwise wait for the copies too.
      if (cq + kStageChannels > tile.rx_base + 8) tile.fill_receive(cq);
      int q = cq;
      for (; q < nq; q = tile.next_live(q)) {
        const Channel ch = setup(q);
        const int2 b = ch.w.at(0);
        if (b.x > b.y) continue;
        const int rows = b.y - b.x + 1;
        if (st.n > 0 && (ch.cfg != st.cfg || rows > max_rows || total + rows > max_rows)) break;
        if (rows > max_rows) {  // first chunk of a long window: a stage of its own
          record(0, q, b.x);
          copy_rows(plane(q) + static_cast<size_t>(b.x) * nt, buf, max_rows);
          cq = q;
          cr = b.x + max_rows - 1;
          return MultiStage{1, ch.cfg, b.x, b.x + max_rows - 2, true};
        }
        record(st.n, q, b.x);
        copy_rows(plane(q) + static_cast<size_t>(b.x) * nt, buf + total * nt, rows);
        total += rows;
        st.n += 1;
        st.cfg = ch.cfg;
        st.valid = true;
        if (st.n == kStageChannels) {
          q = tile.next_live(q);
          break;
        }
      }
      cq = q;
      cr = -1;
      return st;
    };
    auto consume_multi = [&](const MultiStage& st, const float2* base, int s) __attribute__((always_inline)) {
      const int n = uniform(st.n), cfg = uniform(st.cfg), t_lo = uniform(st.t_lo), t_hi = uniform(st.t_hi);
      const int tpos = cfg == 0 ? vy : vx;  // transmit position of this lane's voxel
      const int rpos = cfg == 0 ? vx : vy;  // receive position
      const Term tx = tile.transmit(cfg, 0, tpos);
      for (int k = 0; k < n; ++k) {
        const Term rx = slot_rx[(s * kStageChannels + k) * 8 + rpos];
        const int shift = uniform(slot_shift[s * kStageChannels + k]);
        const Tap tp = tap_of(rx, tx, p.ns);
        if (tp.i0 >= 0 && tp.i0 >= t_lo && tp.i0 <= t_hi) accumulate(tp, static_cast<uint32_t>(tp.i0 + shift), base);
      }
    };
    MultiStage cur{0, 0, -kBig, kBig, false};
    if (cq < nq) cur = plan_issue(stage_a, 0);
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __syncthreads();
    auto step = [&](const float2* now, float2* next, int s_now, int s_next) __attribute__((always_inline)) {
      MultiStage nxt{0, 0, -kBig, kBig, false};
      if (cq < nq) nxt = plan_issue(next, s_next);
      if constexpr (APERTURES) {
        if (!rc_stored && cur.cfg == 1) {  // first CR stage: the RC sums are complete
          store_plane(0);
          rc_stored = true;
        }
      }
      consume_multi(cur, now, s_now);
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
      __syncthreads();
      cur = nxt;
    };
    while (cur.valid) {
      step(stage_a, stage_b, 0, 1);
      if (!cur.valid) break;
      step(stage_b, stage_a, 1, 0);
    }
  } else {
    // ---- stages of all live channels, one barrier per stage, planned by one
    // wave each (wave k % 8 plans stage k + 2) into a ring of descriptors.
    StageDesc* desc = reinterpret_cast<StageDesc*>(lds + desc_offset(p.nc));
    // Plan the stage after prev (first: the tile's first stage) into out.  Whole wave.
    // The stage of windows w that starts at the first live angle at or after
    // a, as plan_stage() packs it, with the angles in lanes: rows per lane,
    // their inclusive prefix sum (DPP), and ballots for the first live angle
    // and the first one that no longer fits.  excl: LDS row offset of each
    // lane's window within the stage.
    auto pack = [&](const Windows& win, int a, Stage& st, int& excl) __attribute__((always_inline)) {
      const int r = (lane >= a && lane < p.na && win.lo <= win.hi) ? win.hi - win.lo + 1 : 0;
      const unsigned long long live = __ballot(r > 0);
      if (live == 0) {
        st.valid = false;
        return;
      }
      const int af = __builtin_ctzll(live);
      const int rf = __builtin_amdgcn_readlane(r, af);
      if (rf > max_rows) {  // first chunk of a window longer than a stage
        st = Stage{af, af + 1, __builtin_amdgcn_readlane(win.lo, af), max_rows, false, true};
        excl = 0;
        return;
      }
      int incl = r;  // lanes 0 .. 15 hold the angles (na <= 16): one DPP row
      incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xf, 0xf, true);  // row_shr:1
      incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xf, 0xf, true);  // row_shr:2
      incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xf, 0xf, true);  // row_shr:4
      incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xf, 0xf, true);  // row_shr:8
      const unsigned long long over = __ballot(incl > max_rows);
      const int a1 = over ? __builtin_ctzll(over) : p.na;
      st = Stage{af, a1, -1, __builtin_amdgcn_readlane(incl, a1 - 1), true, true};
      excl = incl - r;
    };
    // Plan the stage after prev (first: the tile's first stage) into out.  Whole wave.
    auto plan_after = [&](const StageDesc* prev, StageDesc* out, bool first) __attribute__((always_inline)) {
      int q = -1;
      Windows w{kBig, -kBig};
      Stage st{0, 0, -1, 0, true, false};
      int excl = 0;
      if (!first) {
        q = uniform(prev->q);
        if (q < 0) {
          if (lane == 0) out->q = -1;
          return;
        }
        w.lo = prev->lo[lane & 15];
        w.hi = prev->hi[lane & 15];
        if (lane >= 16) w = Windows{kBig, -kBig};
        const Stage cur{uniform(prev->a0), uniform(prev->a1), uniform(prev->r0), uniform(prev->rows),
                        uniform(prev->last) != 0, true};
        if (cur.r0 >= 0 && !cur.last)
          st = next_stage(cur, w, p.na, max_rows);  // next chunk of a long window
        else
          pack(w, cur.a1, st, excl);
      }
      bool fresh = false;
      if (!st.valid) {
        for (int qq = tile.next_live(q); qq < nq; qq = tile.next_live(qq)) {
          w = tile.tile_windows(qq);
          pack(w, 0, st, excl);
          if (st.valid) {
            q = qq;
            fresh = true;
            break;
          }
        }
        if (!st.valid) {
          if (lane == 0) out->q = -1;
          return;
        }
      }
      if (lane < 16) {
        out->lo[lane] = w.lo;
        out->hi[lane] = w.hi;
        out->off[lane] = excl;
        out->base[lane] = (st.r0 < 0 && lane >= st.a0 && lane < st.a1 && w.lo <= w.hi) ? w.lo - excl : kBig;
      }
      if (fresh) {
        const int slot = uniform(8 * (q - tile.rx_base));
        if (lane >= slot && lane < slot + 8) out->rx[lane - slot] = tile.rx_term;
      } else if (lane < 8) {
        out->rx[lane] = prev->rx[lane];
      }
      if (lane == 0) {
        out->q = q;
        out->a0 = st.a0;
        out->a1 = st.a1;
        out->r0 = st.r0;
        out->rows = st.rows;
        out->last = st.last ? 1 : 0;
      }
    };
    // Copies of stage d into buf: whole windows one wave each (window j of the
    // stage by wave (first + j) % 8, so a stage's planning wave, first - 1,
    // copies only its eighth window), a chunk of a long window spread over
    // the waves.
    auto issue_from = [&](const StageDesc* d, float2* buf, int first) __attribute__((always_inline)) {
      const int q = uniform(d->q);
      if (q < 0) return;
      const int cfg = q >= p.nc ? 1 : 0;
      const float2* plane0 = static_cast<const float2*>(cfg == 0 ? p.rc : p.cr) +
                             static_cast<size_t>(q - cfg * p.nc) * p.na * plane_size;
      const int a0 = uniform(d->a0), a1 = uniform(d->a1), r0 = uniform(d->r0);
      if (r0 >= 0) {
        copy_rows(plane0 + a0 * plane_size + static_cast<size_t>(r0) * nt, buf, uniform(d->rows));
        return;
      }
      for (int a = a0 + ((wave - first) & (kVWaves - 1)); a < a1; a += kVWaves) {
        const int lo = uniform(d->lo[a]), hi = uniform(d->hi[a]);
        if (lo > hi) continue;
        copy_rows_wave(plane0 + a * plane_size + static_cast<size_t>(lo) * nt, buf + uniform(d->off[a]) * nt,
                       hi - lo + 1);
      }
    };
    // This wave's voxels, stage d from buf.
    auto consume_from = [&](const StageDesc* d, const float2* base) __attribute__((always_inline)) {
      const int q = uniform(d->q);
      const int cfg = q >= p.nc ? 1 : 0;
      const int a0 = uniform(d->a0), a1 = uniform(d->a1), cr0 = uniform(d->r0);
      const int t_lo = cr0 >= 0 ? cr0 : -kBig;
      const int t_hi = cr0 >= 0 && uniform(d->last) == 0 ? cr0 + max_rows - 2 : kBig;
      const int tpos = cfg == 0 ? vy : vx;  // transmit position of this lane's voxel
      const Term rx = d->rx[cfg == 0 ? vx : vy];
      const int base_l = d->base[lane & 15];
      for (int a = a0; a < a1; ++a) {
        const int rb = cr0 >= 0 ? cr0 : __builtin_amdgcn_readlane(base_l, a);
        if (rb == kBig) continue;
        const Term tx = tile.transmit(cfg, a, tpos);
        int i = rx.i + tx.i;
        float f = rx.f + tx.f;
        const bool carry = f >= 1.0f;
        i += carry ? 1 : 0;
        f = carry ? f - 1.0f : f;
        // A lane-dependent branch around the body (not a `continue`) keeps the
        // angle loop itself wave-uniform.  One predicate, without short-circuit
        // branches: 0 <= i < ns - 1 (or i = ns - 1 with f = 0; inactive terms make
        // i negative) and t_lo <= i <= t_hi, both as unsigned range tests.
        const bool valid = (static_cast<unsigned>(i) < static_cast<unsigned>(p.ns - 1)) |
                           ((i == p.ns - 1) & (f == 0.0f));
        const bool in_chunk = static_cast<unsigned>(i) - static_cast<unsigned>(t_lo) <=
                              static_cast<unsigned>(t_hi) - static_cast<unsigned>(t_lo);
        if (valid & in_chunk) {
          Tap tp;
          tp.w = f;
          // Contractions written out, so that code layout cannot change the rounding.
          tp.c = fmaf(rx.c, tx.c, -(rx.s * tx.s));
          tp.s = fmaf(rx.s, tx.c, rx.c * tx.s);
          accumulate(tp, static_cast<uint32_t>(i - rb), base);
        }
      }
    };

    if (wave == 0) {
      plan_after(nullptr, &desc[0], true);
      plan_after(&desc[0], &desc[1], false);
    }
    __syncthreads();
    issue_from(&desc[0], stage_a, 0);
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __syncthreads();
    int k = 0;  // stage being consumed
    auto step = [&](const float2* now, float2* next) __attribute__((always_inline)) {
      StageDesc* cur = &desc[k % kDescs];
      StageDesc* nxt = &desc[(k + 1) % kDescs];
      issue_from(nxt, next, (k + 1) & (kVWaves - 1));
      if constexpr (APERTURES) {
        if (!rc_stored && uniform(cur->q) >= p.nc) {  // first CR stage: the RC sums are complete
          store_plane(0);
          rc_stored = true;
        }
      }
      consume_from(cur, now);
      // The planning wave plans after its share of the stage: its consumption
      // then starts with the others'.
      if (wave == (k & (kVWaves - 1))) plan_after(nxt, &desc[(k + 2) % kDescs], false);
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
      __syncthreads();
      ++k;
    };
    while (uniform(desc[k % kDescs].q) >= 0) {
      step(stage_a, stage_b);
      if (uniform(desc[k % kDescs].q) < 0) break;
      step(stage_b, stage_a);
    }
  }

  // ---- epilogue: IQ and power (apertures: the last plane, power 0)
  if constexpr (APERTURES) {
    if (!rc_stored) store_plane(0);  // no live CR channel: RC sums, then zeros
    store_plane(1);
    if (tid < kVTileSlots) {
      const long long vv = tile.voxel(tid >> 3, tid & 7);
      if (vv >= 0) {
        p.pd[2 * vv] = 0.f;
        p.pd[2 * vv + 1] = 0.f;
      }
    }
    return;
  }
  float power = 0.f;
#pragma unroll
  for (int k = 0; k < kAcc; ++k) {
    const int t = frame_of(k);
    if (t < nt) {
      float2 s = acc[k];
      if (p.accumulate && v >= 0) {  // complex64 signal holding the sums of other channels
        const float2 prior = reinterpret_cast<const float2*>(p.signal)[v * nt + t];
        s = make_float2(s.x + prior.x, s.y + prior.y);
      }
      power = fmaf(s.x, s.x, fmaf(s.y, s.y, power));
      if (p.signal && v >= 0) {
        const long long index = v * nt + t;
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
  for (int off = 1; off < 8; off <<= 1) power += __shfl_xor(power, off, kWave);
  if (fl == 0) psum[vx * 8 + vy] = power;
  __syncthreads();
  if (tid < kVTileSlots) {
    const long long vv = tile.voxel(tid >> 3, tid & 7);
    if (vv >= 0) {
      p.pd[2 * vv] = psum[tid] / static_cast<float>(nt);
      p.pd[2 * vv + 1] = 0.f;
    }
  }
}

template <int FPP, bool APERTURES, bool TAIL>
hipError_t launch_pairs(GridArgs a) {
  a.max_rows = std::min(kVMaxRows, stage_rows(a.nt, FPP, TAIL));
  if (a.max_rows < kVMinRows) return hipErrorNotSupported;
  const bool multi = a.na == 1;
  const int bytes = dynamic_bytes(a.nc, multi);
  if (bytes + 2 * kVStageWords * 4 > kVLdsBytes) return hipErrorNotSupported;
  const long long grid = (a.tiles + 7) / 8 * 8;
  if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
  // 16-byte rows and pair loads need an even frame count and 16-byte aligned planes.
  if (a.nt % 2 != 0 || reinterpret_cast<uintptr_t>(a.rc) % 16 || reinterpret_cast<uintptr_t>(a.cr) % 16)
    return hipErrorNotSupported;
  a.vec = 1;
  auto kernel =
      multi ? das_opw_grid_valu<FPP, APERTURES, true, TAIL> : das_opw_grid_valu<FPP, APERTURES, false, TAIL>;
  kernel<<<dim3(static_cast<unsigned>(grid)), kVThreads, bytes>>>(a);
  return hipGetLastError();
}

template <bool APERTURES>
hipError_t launch_frames(const GridArgs& a) {
  const int pairs = (a.nt + 15) / 16;
  if (pairs <= 1) return launch_pairs<1, APERTURES, false>(a);
  if (pairs <= 2) return launch_pairs<2, APERTURES, false>(a);
  if (pairs <= 4) return launch_pairs<4, APERTURES, false>(a);
  if (pairs <= 7) return launch_pairs<7, APERTURES, false>(a);
  // At most 8 frames in the last block: single frames there (200 frames: 12 pair blocks + 8).
  if (a.nt <= 16 * (kMaxPairs - 1) + 8) return launch_pairs<kMaxPairs, APERTURES, true>(a);
  return launch_pairs<kMaxPairs, APERTURES, false>(a);
}

}  // namespace

hipError_t launch_opw_grid_valu(const EnsembleArgs& p, const float* x, const float* z, const float* y, int nx,
                                int nz, int ny, int signal_storage, bool apertures, int q_begin, int q_end,
                                bool accumulate) {
  if (p.na > kGridMaxAngles || p.nt > 16 * kMaxPairs) return hipErrorNotSupported;
  if (static_cast<long long>(p.ns) * p.nt > 0x7fffffffLL) return hipErrorNotSupported;
  if (signal_storage < RCABEAM_STORAGE_C64 || signal_storage > RCABEAM_STORAGE_BF16) return hipErrorInvalidValue;
  if (apertures && !p.signal) return hipErrorInvalidValue;  // apertures: complex64 planes whatever signal_storage
  if (accumulate && (apertures || !p.signal || signal_storage != RCABEAM_STORAGE_C64)) return hipErrorInvalidValue;
  GridArgs a = make_grid_args(p, x, z, y, nx, nz, ny, signal_storage);
  if (q_end >= 0) {
    a.q_begin = q_begin;
    a.q_end = q_end;
  }
  a.accumulate = accumulate ? 1 : 0;
  return apertures ? launch_frames<true>(a) : launch_frames<false>(a);
}

}  // namespace rcb
