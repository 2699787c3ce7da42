// Sample-window staging shared by the LDS-tiled delay-and-sum kernels
// (das_tile.hip: VALU interpolation, das_mfma.hip: MFMA interpolation).
//
// A workgroup owns a tile of voxels and, per (aperture, channel), knows each
// angle's exact row window [first row, last row] of the packed (channels,
// angles, samples, frames) layout.  Windows of consecutive angles are packed
// into stages of at most max_rows rows (a longer window is split into chunks
// that overlap by one row); a stage is copied from HBM straight into LDS with
// global_load_lds_dwordx4 (one wave per row, 1 KiB per instruction) while
// the previous stage is being consumed.
#pragma once

#include "common.hpp"

namespace rcb {

constexpr int kBig = 0x7fffffff;

// One complex sample as stored in LDS: float2 for complex64, one 32-bit word
// holding two 16-bit values otherwise.
template <typename T>
struct Word;
template <>
struct Word<float2> {
  using type = float2;
};
template <>
struct Word<__half2> {
  using type = uint32_t;
};
template <>
struct Word<bf16x2> {
  using type = uint32_t;
};

typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf162v __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ float2 widen(typename Word<T>::type w);
template <>
__device__ __forceinline__ float2 widen<float2>(float2 w) {
  return w;
}
template <>
__device__ __forceinline__ float2 widen<__half2>(uint32_t w) {
  const half2v h = __builtin_bit_cast(half2v, w);
  return make_float2(static_cast<float>(h.x), static_cast<float>(h.y));
}
template <>
__device__ __forceinline__ float2 widen<bf16x2>(uint32_t w) {
  return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}

// Two float32 values rounded (to nearest even) into one 32-bit word of the
// 16-bit type: `lo` in the low half.
template <typename T>
__device__ __forceinline__ uint32_t pack2(float lo, float hi);
template <>
__device__ __forceinline__ uint32_t pack2<__half2>(float lo, float hi) {
  half2v h;
  h.x = static_cast<_Float16>(lo);
  h.y = static_cast<_Float16>(hi);
  return __builtin_bit_cast(uint32_t, h);
}
template <>
__device__ __forceinline__ uint32_t pack2<bf16x2>(float lo, float hi) {
  return bf16_pack2_rne(lo, hi);
}

__device__ __forceinline__ int wave_min(int x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = min(x, __shfl_xor(x, off, kWave));
  return x;
}
__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = max(x, __shfl_xor(x, off, kWave));
  return x;
}

// Row windows of one channel held in registers: lane a (< na) holds angle
// a's [first row, last row] (empty: first > last); any lane reads them with
// v_readlane, so stage planning costs no LDS round trips.
struct Windows {
  int lo, hi;
  __device__ __forceinline__ int2 at(int a) const {
    return make_int2(__builtin_amdgcn_readlane(lo, a), __builtin_amdgcn_readlane(hi, a));
  }
};

__device__ __forceinline__ Windows load_windows(const int2* bounds, int na) {
  const int lane = static_cast<int>(threadIdx.x % kWave);
  const int2 b = lane < na ? bounds[lane] : make_int2(kBig, -kBig);
  return Windows{b.x, b.y};
}

// One stage: angles [a0, a1) whose windows are packed back to back, or one
// chunk (rows [r0, r0 + rows)) of a window longer than max_rows (r0 >= 0).
struct Stage {
  int a0, a1, r0, rows;
  bool last, valid;
};

// The stage that starts at the first live angle at or after a.
__device__ __forceinline__ Stage plan_stage(const Windows& w, int a, int na, int max_rows) {
  for (; a < na; ++a) {
    const int2 b = w.at(a);
    if (b.x > b.y) continue;
    const int rows = b.y - b.x + 1;
    if (rows > max_rows) return Stage{a, a + 1, b.x, max_rows, false, true};
    int total = rows, a1 = a + 1;
    for (; a1 < na; ++a1) {
      const int2 b1 = w.at(a1);
      if (b1.x > b1.y) continue;
      const int r1 = b1.y - b1.x + 1;
      if (total + r1 > max_rows) break;
      total += r1;
    }
    return Stage{a, a1, -1, total, true, true};
  }
  return Stage{0, 0, -1, 0, true, false};
}

// plan_stage() with the angles in lanes (na <= 16, so lanes 0 .. 15 form one
// DPP row): rows per lane, their inclusive prefix sum, and ballots for the
// first live angle and the first one that no longer fits, instead of a loop
// of readlanes.  Same stage; *excl (optional) receives each lane's row offset
// within it.  Called by the whole wave.
__device__ __forceinline__ Stage plan_stage_lanes(const Windows& w, int a, int na, int max_rows,
                                                  int* excl = nullptr) {
  const int lane = static_cast<int>(threadIdx.x % kWave);
  const int r = (lane >= a && lane < na && w.lo <= w.hi) ? w.hi - w.lo + 1 : 0;
  const unsigned long long live = __ballot(r > 0);
  if (live == 0) return Stage{0, 0, -1, 0, true, false};
  const int af = __builtin_ctzll(live);
  const int rf = __builtin_amdgcn_readlane(r, af);
  if (rf > max_rows) {  // first chunk of a window longer than a stage
    if (excl) *excl = 0;
    return Stage{af, af + 1, __builtin_amdgcn_readlane(w.lo, af), max_rows, false, true};
  }
  int incl = r;
  incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xf, 0xf, true);  // row_shr:1
  incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xf, 0xf, true);  // row_shr:2
  incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xf, 0xf, true);  // row_shr:4
  incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xf, 0xf, true);  // row_shr:8
  const unsigned long long over = __ballot(incl > max_rows);
  const int a1 = over ? __builtin_ctzll(over) : na;
  if (excl) *excl = incl - r;
  return Stage{af, a1, -1, __builtin_amdgcn_readlane(incl, a1 - 1), true, true};
}

__device__ __forceinline__ Stage next_stage_lanes(const Stage& s, const Windows& w, int na, int max_rows,
                                                  int* excl = nullptr) {
  if (s.r0 >= 0 && !s.last) {
    const int hi = w.at(s.a0).y;
    const int r0 = s.r0 + max_rows - 1;
    const bool last = r0 + max_rows - 1 >= hi;
    return Stage{s.a0, s.a1, r0, last ? hi + 1 - r0 : max_rows, last, true};
  }
  return plan_stage_lanes(w, s.a1, na, max_rows, excl);
}

__device__ __forceinline__ Stage next_stage(const Stage& s, const Windows& w, int na, int max_rows) {
  if (s.r0 >= 0 && !s.last) {
    const int hi = w.at(s.a0).y;
    const int r0 = s.r0 + max_rows - 1;
    const bool last = r0 + max_rows - 1 >= hi;
    return Stage{s.a0, s.a1, r0, last ? hi + 1 - r0 : max_rows, last, true};
  }
  return plan_stage(w, s.a1, na, max_rows);
}

// Copy rows [r_first, r_first + rows) x frames [0, nt) of one plane into LDS
// rows starting at dst.  vec: wave w copies rows w, w + WAVES, ...; each lane
// moves one 16-byte unit of the row straight from HBM into LDS
// (global_load_lds_dwordx4, asynchronous: the caller waits on vmcnt before
// the barrier), so one instruction moves 1 KiB contiguous in both memories.
// Otherwise a synchronous element copy.
template <typename W, int STRIDE, int THREADS>
__device__ __forceinline__ void copy_rows(const W* __restrict__ plane, int nt, int r_first, int rows, W* dst,
                                          bool vec) {
  constexpr int WAVES = THREADS / kWave;
  const W* __restrict__ src = plane + static_cast<size_t>(r_first) * nt;
  if (vec) {
    constexpr int U = 16 / static_cast<int>(sizeof(W));
    const int units = nt / U;  // 16-byte units of one row (nt is a multiple of U)
    const int lane = static_cast<int>(threadIdx.x % kWave);
    for (int row = static_cast<int>(threadIdx.x / kWave); row < rows; row += WAVES) {
      const W* s_row = src + row * nt;
      W* d_row = dst + row * STRIDE;
      for (int c0 = 0; c0 < units; c0 += kWave)
        if (c0 + lane < units)
          __builtin_amdgcn_global_load_lds(
              reinterpret_cast<const void*>(s_row + (c0 + lane) * U),
              reinterpret_cast<__attribute__((address_space(3))) void*>(reinterpret_cast<uintptr_t>(d_row + c0 * U)),
              16, 0, 0);
    }
  } else {
    const int n = rows * STRIDE;
    for (int q = threadIdx.x; q < n; q += THREADS) {
      const int row = q / STRIDE;
      const int col = q - row * STRIDE;
      if (col < nt) dst[q] = src[row * nt + col];
    }
  }
}

template <typename W, int STRIDE, int THREADS>
__device__ __forceinline__ void issue_stage(const W* plane0, size_t plane_size, int nt, const Stage& s,
                                            const Windows& w, W* buf, bool vec) {
  if (s.r0 >= 0) {
    copy_rows<W, STRIDE, THREADS>(plane0 + s.a0 * plane_size, nt, s.r0, s.rows, buf, vec);
    return;
  }
  int off = 0;
  for (int a = s.a0; a < s.a1; ++a) {
    const int2 b = w.at(a);
    if (b.x > b.y) continue;
    copy_rows<W, STRIDE, THREADS>(plane0 + a * plane_size, nt, b.x, b.y - b.x + 1, buf + off * STRIDE, vec);
    off += b.y - b.x + 1;
  }
}

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt unconstrained) in the gfx9 encoding.
constexpr int kWaitVmcnt0 = 0x0F70;

}  // namespace rcb
