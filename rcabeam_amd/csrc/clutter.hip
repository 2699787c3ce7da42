// Temporal SVD clutter filter for MI355X: input scale, Casorati Gram on MFMA,
// FP64 Hermitian eigendecomposition (rocSOLVER), and fused projector +
// filtered power.  Replaces the CuPy/cuBLAS/cuSOLVER path of reference
// src/rcabeam/gpu_clutter.py:112-273 (scaled Gram, eigh, smaller projector,
// float64 residual power).
//
// Gram.  The complex Casorati matrix X (rows x T) is viewed as the real matrix
// A (rows x 2T) with interleaved columns (Re x_t, Im x_t), which is exactly the
// memory layout of complex data.  C = A^T A (2T x 2T, symmetric) gives
//   Re G[t][u] = C[2t][2u] + C[2t+1][2u+1],  Im G[t][u] = C[2t][2u+1] - C[2t+1][2u].
// C is tiled into 32x32 blocks; only the upper-triangular blocks are formed,
// each on v_mfma_f32_32x32x16_{bf16,f16}.  complex64 input is scaled by
// 1/max(|Re|,|Im|) and split into bfloat16 hi + lo (x = hi + lo to ~2^-18
// relative), accumulating all four products hi*hi + hi*lo + lo*hi + lo*lo;
// FP16/BF16 input is exact in one product.  A 16-row step is staged into LDS
// column-major by threads that each convert 8 rows of one frame and write
// them as 16-byte words; two stage buffers alternate (one barrier per step)
// when two workgroups still fit the CU's LDS.  Each workgroup sums a bounded run of rows in float32 (the
// rounding of short runs stays ~1e-8 relative after the cross-workgroup sum)
// and writes a float32 partial; partials are reduced in float64.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <mutex>

#include "common.hpp"
#include "rcabeam_mi355x.h"

namespace rcb {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kGramThreads = 512;             // 8 waves, 2 per SIMD
constexpr int kGramWaves = kGramThreads / kWave;
constexpr int kGramMaxBlocksPerWave = 12;
constexpr int kGramBlocksPerPass = kGramWaves * kGramMaxBlocksPerWave;  // 96
constexpr int kGramK = 16;                    // rows per MFMA step
constexpr int kGramKPad = kGramK + 8;         // LDS column stride (bf16 elements, 48 bytes)
constexpr int kGramRowsPerGroup = 2048;       // rows summed in float32 by one workgroup
constexpr int kGramMaxColumnBlocks = 64;      // 2T <= 2048

struct GramPlan {
  int nb;                                     // column blocks of 32 (2T padded)
  int wave_start[kGramWaves];
  int wave_count[kGramWaves];
  unsigned char bi[kGramBlocksPerPass];
  unsigned char bj[kGramBlocksPerPass];
  int pass_first;                             // index of this pass's first block in the full list
  int pass_blocks;
};

enum GramMode { kSplitBf16 = 0, kExactF16 = 1, kExactBf16 = 2 };

// MODE kSplitBf16: T = float2 (complex64), hi/lo bfloat16.
// MODE kExactF16:  T = __half2, one fp16 product.
// MODE kExactBf16: T = bf16x2, one bf16 product.
// Row r of the Gram is row r * row_step of x (row_step 1: every row).
template <int MODE, bool DB, typename T>
__global__ __launch_bounds__(kGramThreads) void k_gram(const T* __restrict__ x, long long rows, long long row_step,
                                                      int frames, const float* __restrict__ scale_ptr, GramPlan plan,
                                                      float* __restrict__ partials) {
  constexpr bool kSplit = MODE == kSplitBf16;
  // [buffer (DB: 2)][hi, lo][cols][kGramKPad] bf16 / f16 operands of one 16-row step, column-major.
  extern __shared__ __align__(16) unsigned short opnd[];
  const int cols = plan.nb * 32;
  const int buf_elems = 2 * cols * kGramKPad;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  // 1 / max(|Re|, |Im|) in float64: the max may be a float32 subnormal whose inverse overflows float32.
  const float maxabs = kSplit ? *scale_ptr : 1.0f;
  const double inv_scale = maxabs > 0.f ? 1.0 / static_cast<double>(maxabs) : 1.0;
  const long long row0 = static_cast<long long>(blockIdx.x) * kGramRowsPerGroup;
  const long long row_end = row0 + kGramRowsPerGroup < rows ? row0 + kGramRowsPerGroup : rows;
  const int steps = static_cast<int>((row_end - row0 + kGramK - 1) / kGramK);

  // Zero the padding columns once (they stay zero).
  for (int b = 0; b < (DB ? 2 : 1); ++b)
    for (int idx = threadIdx.x; idx < (cols - 2 * frames) * kGramKPad; idx += blockDim.x) {
      opnd[b * buf_elems + 2 * frames * kGramKPad + idx] = 0;
      opnd[b * buf_elems + cols * kGramKPad + 2 * frames * kGramKPad + idx] = 0;
    }

  // Stage step s into buffer base: thread item (h, t) converts rows 8 h .. 8 h + 7 of frame t and
  // writes them as one 16-byte word per operand column (Re and Im of hi, and of lo when split).
  auto stage = [&](int s, unsigned short* base) __attribute__((always_inline)) {
    unsigned short* hi = base;
    unsigned short* lo = base + cols * kGramKPad;
    const long long rbase = row0 + static_cast<long long>(s) * kGramK;
    for (int item = threadIdx.x; item < 2 * frames; item += kGramThreads) {
      const int h = item / frames, t = item - h * frames;
      float2 v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const long long r = rbase + 8 * h + j;
        v[j] = r < row_end ? load_sample(x + r * row_step * frames + t) : make_float2(0.f, 0.f);
      }
      uint32_t hre[4] = {0u, 0u, 0u, 0u}, him[4] = {0u, 0u, 0u, 0u}, lre[4] = {0u, 0u, 0u, 0u}, lim[4] = {0u, 0u, 0u, 0u};
      if constexpr (MODE == kExactF16) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int sh = 16 * (j & 1);
          hre[j >> 1] |= static_cast<uint32_t>(__half_as_ushort(__float2half_rn(v[j].x))) << sh;
          him[j >> 1] |= static_cast<uint32_t>(__half_as_ushort(__float2half_rn(v[j].y))) << sh;
        }
      } else {
        // Rows 2 m and 2 m + 1 of a column share a word: one v_cvt_pk_bf16_f32 per word (round to
        // nearest even, as the integer form), and the low parts from the words' high halves.
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const float x0 = static_cast<float>(static_cast<double>(v[2 * m].x) * inv_scale);
          const float x1 = static_cast<float>(static_cast<double>(v[2 * m + 1].x) * inv_scale);
          const float y0 = static_cast<float>(static_cast<double>(v[2 * m].y) * inv_scale);
          const float y1 = static_cast<float>(static_cast<double>(v[2 * m + 1].y) * inv_scale);
          hre[m] = bf16_pack2_rne(x0, x1);
          him[m] = bf16_pack2_rne(y0, y1);
          if constexpr (kSplit) {
            lre[m] = bf16_pack2_rne(x0 - __uint_as_float(hre[m] << 16), x1 - __uint_as_float(hre[m] & 0xffff0000u));
            lim[m] = bf16_pack2_rne(y0 - __uint_as_float(him[m] << 16), y1 - __uint_as_float(him[m] & 0xffff0000u));
          }
        }
      }
      *reinterpret_cast<uint4*>(hi + (2 * t) * kGramKPad + 8 * h) = make_uint4(hre[0], hre[1], hre[2], hre[3]);
      *reinterpret_cast<uint4*>(hi + (2 * t + 1) * kGramKPad + 8 * h) = make_uint4(him[0], him[1], him[2], him[3]);
      if constexpr (kSplit) {
        *reinterpret_cast<uint4*>(lo + (2 * t) * kGramKPad + 8 * h) = make_uint4(lre[0], lre[1], lre[2], lre[3]);
        *reinterpret_cast<uint4*>(lo + (2 * t + 1) * kGramKPad + 8 * h) = make_uint4(lim[0], lim[1], lim[2], lim[3]);
      }
    }
  };

  floatx16 acc[kGramMaxBlocksPerWave];
#pragma unroll
  for (int b = 0; b < kGramMaxBlocksPerWave; ++b)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[b][e] = 0.f;

  const int my_start = plan.wave_start[wave];
  const int my_count = plan.wave_count[wave];

  // This wave's blocks of one staged step.
  auto compute = [&](const unsigned short* base) __attribute__((always_inline)) {
    const unsigned short* hi = base;
    const unsigned short* lo = base + cols * kGramKPad;
    const int koff = 8 * (lane / 32);
    int prev = -1;
    bf16x8 a_hi, a_lo;
#pragma unroll
    for (int b = 0; b < kGramMaxBlocksPerWave; ++b) {
      if (b < my_count) {
        const int I = plan.bi[my_start + b];
        const int J = plan.bj[my_start + b];
        if (I != prev) {
          a_hi = *reinterpret_cast<const bf16x8*>(hi + (32 * I + (lane & 31)) * kGramKPad + koff);
          if constexpr (kSplit) a_lo = *reinterpret_cast<const bf16x8*>(lo + (32 * I + (lane & 31)) * kGramKPad + koff);
          prev = I;
        }
        const bf16x8 b_hi = *reinterpret_cast<const bf16x8*>(hi + (32 * J + (lane & 31)) * kGramKPad + koff);
        if constexpr (MODE == kExactF16) {
          acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a_hi),
                                                          __builtin_bit_cast(f16x8, b_hi), acc[b], 0, 0, 0);
        } else {
          acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_hi, b_hi, acc[b], 0, 0, 0);
          if constexpr (kSplit) {
            const bf16x8 b_lo = *reinterpret_cast<const bf16x8*>(lo + (32 * J + (lane & 31)) * kGramKPad + koff);
            acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_hi, b_lo, acc[b], 0, 0, 0);
            acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_lo, b_hi, acc[b], 0, 0, 0);
            // lo * lo too: dropping it biases the diagonal by -sum lo^2 (always one sign).
            acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_lo, b_lo, acc[b], 0, 0, 0);
          }
        }
      }
    }
  };

  if constexpr (DB) {
    // Double-buffered: step s + 1 is staged while step s is multiplied, one barrier per step.
    if (steps > 0) stage(0, opnd);
    __syncthreads();
    for (int s = 0; s < steps; ++s) {
      if (s + 1 < steps) stage(s + 1, opnd + ((s + 1) & 1) * buf_elems);
      compute(opnd + (s & 1) * buf_elems);
      __syncthreads();
    }
  } else {
    for (int s = 0; s < steps; ++s) {
      __syncthreads();
      stage(s, opnd);
      __syncthreads();
      compute(opnd);
    }
  }
  // Partials: [group][pass block][lane][16]
  float* out = partials + (static_cast<long long>(blockIdx.x) * plan.pass_blocks) * 1024;
#pragma unroll
  for (int b = 0; b < kGramMaxBlocksPerWave; ++b) {
    if (b < my_count) {
      float* dst = out + static_cast<long long>(my_start + b) * 1024 + lane * 16;
#pragma unroll
      for (int e = 0; e < 16; ++e) dst[e] = acc[b][e];
    }
  }
}

// Sum the float32 partials of all row groups in float64 into the real Gram
// C (ldc = 32 nb, row-major, upper blocks only), optionally accumulating.
__global__ void k_gram_reduce(const float* __restrict__ partials, int groups, GramPlan plan,
                              double* __restrict__ c, int accumulate) {
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long long>(plan.pass_blocks) * 1024) return;
  const int b = static_cast<int>(idx / 1024);
  const int e = static_cast<int>(idx % 1024);
  double sum = 0.0;
  for (int g = 0; g < groups; ++g) sum += static_cast<double>(partials[(static_cast<long long>(g) * plan.pass_blocks + b) * 1024 + e]);
  const int l = e / 16, r = e % 16;
  // 32x32 MFMA accumulator layout: lane l, register r -> (row, col).
  const int i = 8 * (r / 4) + 4 * (l / 32) + (r % 4);
  const int j = l % 32;
  const int I = plan.bi[b], J = plan.bj[b];
  const long long ldc = 32LL * plan.nb;
  double* dst = c + (32LL * I + i) * ldc + 32LL * J + j;
  *dst = accumulate ? *dst + sum : sum;
}

// Upper-triangular blocks are stored (diagonal blocks in full), so any entry
// with i <= j is present and C is symmetric.
__device__ __forceinline__ double c_at(const double* c, long long ldc, int i, int j) {
  return i <= j ? c[i * ldc + j] : c[j * ldc + i];
}

// G (complex128, column-major T x T) = scale2 * complex Gram from C.
__global__ void k_gram_form(const double* __restrict__ c, int nb, int frames, double scale2, double2* __restrict__ g) {
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long long>(frames) * frames) return;
  const int t = static_cast<int>(idx % frames);  // row
  const int u = static_cast<int>(idx / frames);  // column
  const long long ldc = 32LL * nb;
  const double re = c_at(c, ldc, 2 * t, 2 * u) + c_at(c, ldc, 2 * t + 1, 2 * u + 1);
  const double im = c_at(c, ldc, 2 * t, 2 * u + 1) - c_at(c, ldc, 2 * t + 1, 2 * u);
  g[idx] = make_double2(re * scale2, im * scale2);
}

// ---- exact Gram: float64 MFMA ----------------------------------------------
// complex64 samples widen exactly to float64; v_mfma_f64_16x16x4_f64 forms the
// products exactly and accumulates in float64, so C equals the complex128
// reference Gram up to float64 summation order.  C is tiled into 16x16 blocks
// and its upper-triangular blocks (row-major order) are cut into sets of up to
// kG64BlocksPerSet; rows are cut into groups.  One launch covers every (group,
// set) pair: workgroup (g, s) stages the rows of group g in LDS 16 at a time
// and forms the blocks of set s.  The sets of one group are launched next to
// each other on the same XCD (workgroup w runs on XCD w % 8), so a row chunk
// is fetched from HBM once and served to the other sets from that XCD's L2.
// Each thread prefetches its share of the next step's rows into registers
// while the waves run the MFMAs of the current ones.  A step is 32 rows up to
// 224 frames (fewer barriers: 14.5 against 15.4 ms at 4,194,304 x 200) and 16
// rows beyond, where 32 would not fit in LDS.  Partials are [group][block]
// [256] float64 and are summed over groups in a fixed order (deterministic).
typedef double doublex4 __attribute__((ext_vector_type(4)));
constexpr int kG64Threads = 1024;
constexpr int kG64Waves = kG64Threads / kWave;
constexpr int kG64MaxBlocks = 7;
constexpr int kG64BlocksPerSet = kG64Waves * kG64MaxBlocks;  // 112
constexpr int kG64MaxFrames = 640;
constexpr int kG64ShortFrames = 224;                         // frames of the 32-row step form
constexpr int kG64MaxRows = 32;                              // the longest step
constexpr long long kG64RowsPerGroup = 16384;                // default grouping (multiple of kG64MaxRows)
__host__ __device__ constexpr int g64_prefetch(int rows, int max_frames) {  // float2 per thread
  return (rows * max_frames + kG64Threads - 1) / kG64Threads;
}
inline int g64_step_rows(size_t frames) { return frames <= static_cast<size_t>(kG64ShortFrames) ? 32 : 16; }

struct Gram64Shape {
  int nb;                   // 16-column blocks of the real Gram (2T padded to 16)
  int total;                // upper-triangular blocks
  int sets;                 // ceil(total / kG64BlocksPerSet)
  int groups;               // row groups
  long long rows_per_group; // multiple of kG64MaxRows
};

// (I, J) of upper-triangular block `idx` in row-major order of an nb x nb block grid.
__device__ __forceinline__ int2 triangle_block(int idx, int nb) {
  int i = 0;
  while (i < nb - 1 && idx >= nb - i) {
    idx -= nb - i;
    ++i;
  }
  return make_int2(i, i + idx);
}

template <int ROWS, int MAXF>
__global__ __launch_bounds__(kG64Threads) void k_gram64(const float2* __restrict__ x, long long rows, int frames,
                                                       Gram64Shape shape, double* __restrict__ partials) {
  constexpr int kPrefetch = g64_prefetch(ROWS, MAXF);
  extern __shared__ __align__(16) double stage[];  // [ROWS][ld]
  // Workgroup -> (group, set): the sets of 8 consecutive groups occupy 8 * sets consecutive workgroups,
  // set-major, so group g's sets all land on XCD g % 8 at about the same time.
  const int per_chunk = 8 * shape.sets;
  const int chunk = static_cast<int>(blockIdx.x) / per_chunk, rem = static_cast<int>(blockIdx.x) % per_chunk;
  const int set = rem / 8;
  const int group = chunk * 8 + rem % 8;
  if (group >= shape.groups) return;
  const int ld = shape.nb * 16;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = uniform(static_cast<int>(threadIdx.x) / kWave);
  const long long row0 = static_cast<long long>(group) * shape.rows_per_group;
  const long long row_end = row0 + shape.rows_per_group < rows ? row0 + shape.rows_per_group : rows;
  const int steps = static_cast<int>((row_end - row0 + ROWS - 1) / ROWS);
  // This wave's run of the set's blocks.
  const int first = set * kG64BlocksPerSet;
  const int count = shape.total - first < kG64BlocksPerSet ? shape.total - first : kG64BlocksPerSet;
  const int base = count / kG64Waves, extra = count % kG64Waves;
  const int my_count = base + (wave < extra ? 1 : 0);
  const int my_start = first + wave * base + (wave < extra ? wave : extra);
  // Stage offsets of the two 16-column slices each block multiplies.  Blocks
  // past my_count read slice 0 and are computed but never stored, so every
  // wave runs one straight-line sequence per 4 rows: all operand reads, then
  // the MFMAs back to back (no LDS wait between them).
  int aoff[kG64MaxBlocks], boff[kG64MaxBlocks];
  {
    int2 ij = triangle_block(my_start, shape.nb);
#pragma unroll
    for (int b = 0; b < kG64MaxBlocks; ++b) {
      aoff[b] = b < my_count ? 16 * ij.x : 0;
      boff[b] = b < my_count ? 16 * ij.y : 0;
      if (++ij.y == shape.nb) {
        ++ij.x;
        ij.y = ij.x;
      }
    }
  }
  for (int idx = threadIdx.x; idx < ROWS * (ld - 2 * frames); idx += blockDim.x) {
    const int k = idx / (ld - 2 * frames), c = idx % (ld - 2 * frames);
    stage[k * ld + 2 * frames + c] = 0.0;
  }
  doublex4 acc[kG64MaxBlocks];
#pragma unroll
  for (int b = 0; b < kG64MaxBlocks; ++b) acc[b] = doublex4{0.0, 0.0, 0.0, 0.0};
  const int per_step = ROWS * frames;
  // Prefetched share of the next step: element idx = u * kG64Threads + tid of the step's ROWS x frames samples.
  float2 next[kPrefetch];
  auto fetch = [&](int s) {
    const long long rbase = row0 + static_cast<long long>(s) * ROWS;
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
      const int idx = u * kG64Threads + static_cast<int>(threadIdx.x);
      next[u] = make_float2(0.f, 0.f);
      if (idx < per_step) {
        const int k = idx / frames;
        if (rbase + k < row_end) next[u] = x[rbase * frames + idx];
      }
    }
  };
  if (steps > 0) fetch(0);
  for (int s = 0; s < steps; ++s) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
      const int idx = u * kG64Threads + static_cast<int>(threadIdx.x);
      if (idx < per_step) {
        const int k = idx / frames, t = idx - k * frames;
        *reinterpret_cast<double2*>(stage + k * ld + 2 * t) = make_double2(next[u].x, next[u].y);
      }
    }
    __syncthreads();
    if (s + 1 < steps) fetch(s + 1);
    // This lane's first stage element, opaque to the compiler, so that the
    // 14 read addresses of each 4-row block are formed where they are used
    // instead of all being hoisted out of the step loop (and spilled).
    int lane_elem = (lane / 16) * ld + (lane & 15);
    asm volatile("" : "+v"(lane_elem));
    // Operand reads in flight per group of blocks: all 7 where the prefetch
    // leaves room in the 128 registers, 4 then 3 otherwise.
    constexpr int kGroup = kPrefetch > 8 ? 4 : kG64MaxBlocks;
#pragma unroll
    for (int kk = 0; kk < ROWS; kk += 4) {
      const double* rowp = stage + kk * ld + lane_elem;
#pragma unroll
      for (int b0 = 0; b0 < kG64MaxBlocks; b0 += kGroup) {
        double a[kGroup], bv[kGroup];
#pragma unroll
        for (int b = 0; b < kGroup; ++b)
          if (b0 + b < kG64MaxBlocks) {
            a[b] = rowp[aoff[b0 + b]];
            bv[b] = rowp[boff[b0 + b]];
          }
#pragma unroll
        for (int b = 0; b < kGroup; ++b)
          if (b0 + b < kG64MaxBlocks)
            acc[b0 + b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[b], bv[b], acc[b0 + b], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);  // keep the next reads after these MFMAs
      }
    }
  }
  double* out = partials + (static_cast<long long>(group) * shape.total + my_start) * 256;
#pragma unroll
  for (int b = 0; b < kG64MaxBlocks; ++b)
    if (b < my_count) {
      double* dst = out + static_cast<long long>(b) * 256 + lane * 4;
#pragma unroll
      for (int e = 0; e < 4; ++e) dst[e] = acc[b][e];
    }
}

// Sum float64 partials over groups (in group order) into the real Gram C (ldc = 32 * ceil(2T / 32)).
__global__ void k_gram64_reduce(const double* __restrict__ partials, Gram64Shape shape, int ldc, double* __restrict__ c,
                                int accumulate) {
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long long>(shape.total) * 256) return;
  const int b = static_cast<int>(idx / 256), e = static_cast<int>(idx % 256);
  double sum = 0.0;
  for (int g = 0; g < shape.groups; ++g) sum += partials[(static_cast<long long>(g) * shape.total + b) * 256 + e];
  const int l = e / 4, r = e % 4;
  const int2 ij = triangle_block(b, shape.nb);
  // 16x16x4 float64 MFMA accumulator layout (measured on gfx950): lane l,
  // register r -> (row l/16 + 4 r, column l%16).
  const int i = 16 * ij.x + (l / 16) + 4 * r;
  const int j = 16 * ij.y + (l % 16);
  double* dst = c + static_cast<long long>(i) * ldc + j;
  *dst = accumulate ? *dst + sum : sum;
}

// ---- max(|Re|, |Im|) and finiteness ---------------------------------------
__global__ void k_zero2(unsigned* out) {
  if (threadIdx.x < 3) out[threadIdx.x] = 0u;
}

// out[2] = bits of 1 / max (1 when the input is all zero).
__global__ void k_inverse_scale(unsigned* out) {
  if (threadIdx.x == 0) {
    const float m = __uint_as_float(out[0]);
    out[2] = __float_as_uint(m > 0.f ? 1.0f / m : 1.0f);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_absmax(const T* __restrict__ x, long long n, unsigned* out) {
  float m = 0.f;
  bool bad = false;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const float2 v = load_sample(x + i);
    bad |= !isfinite(v.x) || !isfinite(v.y);
    m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, kWave));
  const bool any_bad = __any(bad);
  if ((threadIdx.x & (kWave - 1)) == 0) {
    atomicMax(out, __float_as_uint(m));
    if (any_bad) atomicOr(out + 1, 1u);
  }
}

// ---- projector + filtered power ---------------------------------------------
// One wave per Casorati row x (a row vector).  Coefficients c = x V, i.e.
// c_k = sum_t x_t V[t,k], then x - c V^H (subtract form) or |c|^2 (retain
// form), as block @ vectors @ vectors^H in reference gpu_clutter.py:215.
// The residual / retained energy are formed in float64 (the operation is
// HBM-bound; float64 removes the cancellation of x - P x for strong clutter).
template <typename T, bool FILTER, int kMaxPerLane>
__global__ __launch_bounds__(256) void k_project(const T* __restrict__ x, long long rows, int frames,
                                                 const double2* __restrict__ v, int r, int subtract,
                                                 double* __restrict__ power, float2* __restrict__ filtered) {
  const long long row = static_cast<long long>(blockIdx.x) * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= rows) return;
  const T* xr = x + row * frames;
  double2 xv[kMaxPerLane];
#pragma unroll
  for (int j = 0; j < kMaxPerLane; ++j) {
    const int t = lane + kWave * j;
    const float2 s = t < frames ? load_sample(xr + t) : make_float2(0.f, 0.f);
    xv[j] = make_double2(s.x, s.y);
  }
  double2 resid[kMaxPerLane];
#pragma unroll
  for (int j = 0; j < kMaxPerLane; ++j) resid[j] = subtract ? xv[j] : make_double2(0.0, 0.0);
  double kept = 0.0;
  for (int k = 0; k < r; ++k) {
    const double2* vk = v + static_cast<long long>(k) * frames;
    double cr = 0.0, ci = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxPerLane; ++j) {
      const int t = lane + kWave * j;
      if (t < frames) {
        const double2 b = vk[t];
        cr += xv[j].x * b.x - xv[j].y * b.y;  // c_k = sum_t x_t V[t, k]
        ci += xv[j].x * b.y + xv[j].y * b.x;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      cr += __shfl_xor(cr, off, kWave);
      ci += __shfl_xor(ci, off, kWave);
    }
    if (!subtract && !FILTER) {
      kept += cr * cr + ci * ci;
      continue;
    }
    const double sign = subtract ? -1.0 : 1.0;
#pragma unroll
    for (int j = 0; j < kMaxPerLane; ++j) {
      const int t = lane + kWave * j;
      if (t < frames) {
        const double2 b = vk[t];  // projection c_k conj(V[t, k]) (row-vector convention x V V^H)
        resid[j].x += sign * (cr * b.x + ci * b.y);
        resid[j].y += sign * (ci * b.x - cr * b.y);
      }
    }
  }
  if constexpr (FILTER) {
#pragma unroll
    for (int j = 0; j < kMaxPerLane; ++j) {
      const int t = lane + kWave * j;
      if (t < frames) filtered[row * frames + t] = make_float2(static_cast<float>(resid[j].x), static_cast<float>(resid[j].y));
    }
  } else {
    double e = 0.0;
    if (subtract || FILTER) {
#pragma unroll
      for (int j = 0; j < kMaxPerLane; ++j) e += resid[j].x * resid[j].x + resid[j].y * resid[j].y;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, kWave);
    } else {
      e = kept;
    }
    if (lane == 0) power[row] = e / frames;
  }
}

// The same projection for short rows (frames <= 256) and few vectors (r <= 8),
// the clutter filter's usual shape: L lanes per row (lane f holds frames
// f + L j, j < kPer; 8 lanes up to 128 frames, 16 beyond, so a load reads 64
// or 128 contiguous bytes of a row), 64 / L rows per wave, so the coordinate
// reductions stay inside the row's lanes; the vectors are staged in LDS once
// per workgroup.  The residual x_t - sum_k c_k conj(V[t, k]) is formed frame
// by frame in float64 from the float32 row kept in registers.  R bounds r at
// compile time (registers for R coordinates only).
constexpr int kProjectMaxR = 8;

template <typename T, bool FILTER, int L, int kPer, int R>
__global__ __launch_bounds__(256) void k_project_rows8(const T* __restrict__ x, long long rows, int frames,
                                                       const double2* __restrict__ v, int r, int subtract,
                                                       double* __restrict__ power, float2* __restrict__ filtered) {
  extern __shared__ double2 vs[];  // [r][frames]
  for (int i = threadIdx.x; i < r * frames; i += blockDim.x) vs[i] = v[i];
  __syncthreads();
  const long long row = static_cast<long long>(blockIdx.x) * (256 / L) + threadIdx.x / L;
  const int f = threadIdx.x & (L - 1);
  if (row >= rows) return;  // after the only barrier
  const T* xr = x + row * frames;
  float2 xv[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int t = f + L * j;
    xv[j] = t < frames ? load_sample(xr + t) : make_float2(0.f, 0.f);
  }
  // Coordinates c_k = sum_t x_t V[t, k] of this row (all L lanes hold them after the reduction).
  double cre[R], cim[R];
  double kept = 0.0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    cre[k] = cim[k] = 0.0;
    if (k < r) {
      const double2* vk = vs + k * frames;
#pragma unroll
      for (int j = 0; j < kPer; ++j) {
        const int t = f + L * j;
        if (t < frames) {
          const double2 b = vk[t];
          const double xr_ = xv[j].x, xi_ = xv[j].y;
          cre[k] += xr_ * b.x - xi_ * b.y;
          cim[k] += xr_ * b.y + xi_ * b.x;
        }
      }
#pragma unroll
      for (int off = 1; off < L; off <<= 1) {
        cre[k] += __shfl_xor(cre[k], off, kWave);
        cim[k] += __shfl_xor(cim[k], off, kWave);
      }
      kept += cre[k] * cre[k] + cim[k] * cim[k];
    }
  }
  if (!FILTER && !subtract) {  // retained energy |c|^2 directly
    if (f == 0) power[row] = kept / frames;
    return;
  }
  const double sign = subtract ? -1.0 : 1.0;
  double e = 0.0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int t = f + L * j;
    if (t < frames) {
      double re = subtract ? static_cast<double>(xv[j].x) : 0.0;
      double im = subtract ? static_cast<double>(xv[j].y) : 0.0;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        if (k < r) {
          const double2 b = vs[k * frames + t];  // c_k conj(V[t, k]) (row-vector convention x V V^H)
          re += sign * (cre[k] * b.x + cim[k] * b.y);
          im += sign * (cim[k] * b.x - cre[k] * b.y);
        }
      }
      if constexpr (FILTER)
        filtered[row * frames + t] = make_float2(static_cast<float>(re), static_cast<float>(im));
      else
        e += re * re + im * im;
    }
  }
  if constexpr (!FILTER) {
#pragma unroll
    for (int off = 1; off < L; off <<= 1) e += __shfl_xor(e, off, kWave);
    if (f == 0) power[row] = e / frames;
  }
}

// ---- host helpers -----------------------------------------------------------

static int column_blocks(size_t frames) { return static_cast<int>((2 * frames + 31) / 32); }

static long long triangle_blocks(int nb) { return static_cast<long long>(nb) * (nb + 1) / 2; }

static void make_plan(int nb, int pass, GramPlan& plan) {
  plan.nb = nb;
  const long long total = triangle_blocks(nb);
  const long long first = static_cast<long long>(pass) * kGramBlocksPerPass;
  const int count = static_cast<int>(total - first < kGramBlocksPerPass ? total - first : kGramBlocksPerPass);
  plan.pass_first = static_cast<int>(first);
  plan.pass_blocks = count;
  long long idx = 0;
  int filled = 0;
  for (int I = 0; I < nb; ++I)
    for (int J = I; J < nb; ++J, ++idx)
      if (idx >= first && idx < first + count) {
        plan.bi[filled] = static_cast<unsigned char>(I);
        plan.bj[filled] = static_cast<unsigned char>(J);
        ++filled;
      }
  // Contiguous runs of the row-major triangle: consecutive blocks mostly share I.
  const int base = count / kGramWaves, extra = count % kGramWaves;
  int start = 0;
  for (int w = 0; w < kGramWaves; ++w) {
    const int n = base + (w < extra ? 1 : 0);
    plan.wave_start[w] = start;
    plan.wave_count[w] = n;
    start += n;
  }
}

static size_t lds_bytes(int nb) { return static_cast<size_t>(2) * nb * 32 * kGramKPad * sizeof(unsigned short); }

template <int MODE, typename T>
static hipError_t gram_launch(const void* x, long long rows, long long row_step, int frames, const float* scale,
                              double* c, float* workspace, int accumulate) {
  const int nb = column_blocks(frames);
  const long long groups = (rows + kGramRowsPerGroup - 1) / kGramRowsPerGroup;
  const int passes = static_cast<int>((triangle_blocks(nb) + kGramBlocksPerPass - 1) / kGramBlocksPerPass);
  // Double-buffered staging when two workgroups per CU still fit in 160 KiB of LDS.
  const bool db = 4 * lds_bytes(nb) <= 160 * 1024;
  const size_t lds = (db ? 2 : 1) * lds_bytes(nb);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  auto kernel = db ? &k_gram<MODE, true, T> : &k_gram<MODE, false, T>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     static_cast<int>(lds));
  if (e != hipSuccess) return e;
  for (int pass = 0; pass < passes; ++pass) {
    GramPlan plan;
    make_plan(nb, pass, plan);
    kernel<<<dim3(static_cast<unsigned>(groups)), kGramThreads, lds>>>(static_cast<const T*>(x), rows, row_step, frames,
                                                                        scale, plan, workspace);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const long long n = static_cast<long long>(plan.pass_blocks) * 1024;
    k_gram_reduce<<<dim3(static_cast<unsigned>((n + 255) / 256)), 256>>>(workspace, static_cast<int>(groups), plan, c,
                                                                          accumulate);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

static int column_blocks16(size_t frames) { return static_cast<int>((2 * frames + 15) / 16); }
static size_t lds64_bytes(size_t frames) {
  return static_cast<size_t>(g64_step_rows(frames)) * 16 * column_blocks16(frames) * sizeof(double);
}

// Shape of the exact Gram for `rows` x `frames` with groups of `rows_per_group` rows.
static Gram64Shape gram64_shape(size_t rows, size_t frames, long long rows_per_group) {
  Gram64Shape sh;
  sh.nb = column_blocks16(frames);
  sh.total = static_cast<int>(triangle_blocks(sh.nb));
  sh.sets = (sh.total + kG64BlocksPerSet - 1) / kG64BlocksPerSet;
  sh.rows_per_group = rows_per_group;
  sh.groups = static_cast<int>((static_cast<long long>(rows) + rows_per_group - 1) / rows_per_group);
  return sh;
}

static size_t gram64_partial_bytes(const Gram64Shape& sh) {
  return static_cast<size_t>(sh.groups) * sh.total * 256 * sizeof(double);
}

static bool gram64_frames_ok(size_t frames) { return frames >= 1 && frames <= kG64MaxFrames; }

// One launch over all (group, set) pairs, then the fixed-order reduction into C.
static hipError_t gram64_launch(const float2* iq, size_t rows, size_t frames, const Gram64Shape& sh,
                                double* partials, double* real_gram, int accumulate) {
  const size_t lds = lds64_bytes(frames);
  auto kernel = g64_step_rows(frames) == 32 ? &k_gram64<32, kG64ShortFrames> : &k_gram64<16, kG64MaxFrames>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
  if (e != hipSuccess) return e;
  const long long chunks = (sh.groups + 7) / 8;
  const long long grid = chunks * 8 * sh.sets;
  if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
  kernel<<<dim3(static_cast<unsigned>(grid)), kG64Threads, lds>>>(iq, static_cast<long long>(rows),
                                                                    static_cast<int>(frames), sh, partials);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const long long n = static_cast<long long>(sh.total) * 256;
  const int ldc = 32 * column_blocks(frames);
  k_gram64_reduce<<<dim3(static_cast<unsigned>((n + 255) / 256)), 256>>>(partials, sh, ldc, real_gram, accumulate);
  return hipGetLastError();
}

// ---- Hermitian EVD, T <= kEvdMaxN: one-workgroup tridiagonalization ---------
// rocSOLVER's zheevd at T = 200 spends 3.7 of its 5.6 ms in zhetrd, ~800
// launches of a few microseconds each.  k_hetrd_q does the whole reduction in
// one workgroup (the matrix stays in L2): Householder reflectors
// H_i = I - tau_i u_i u_i^H (u_i[i + 1] = 1, u_i[i + 2 ..] below the
// subdiagonal) take the lower triangle to a real tridiagonal T = Q^H A Q with
// Q = H_0 H_1 ... H_{n-2} (the steps of LAPACK's zhetd2 'L' with zlarfg's
// choice of beta and tau); then Q is accumulated in place from the last
// reflector to the first (as zung2l/zungtr do), so rocSOLVER's zstedc
// ('original' vectors) returns the eigenvectors of A.  The input's upper
// triangle is the one referenced (the lower is rebuilt from it first), as
// with zheevd(fill_upper).
constexpr int kEvdMaxN = 256;
constexpr int kEvdThreads = 1024;
constexpr int kEvdWaves = kEvdThreads / kWave;

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zmulc(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// The lower triangle from the upper one and a real diagonal (rocSOLVER's
// zheevd reads both, and k_hetrd_q does the same in its prologue).
__global__ void k_hermitian_from_upper(double2* __restrict__ a, int n) {
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long long>(n) * n) return;
  const int r = static_cast<int>(idx % n), c = static_cast<int>(idx / n);
  if (r > c) {
    const double2 u = a[c + static_cast<size_t>(r) * n];
    a[idx] = make_double2(u.x, -u.y);
  } else if (r == c) {
    a[idx].y = 0.0;
  }
}

__global__ __launch_bounds__(kEvdThreads) void k_hetrd_q(double2* __restrict__ a, int n, double* __restrict__ d,
                                                        double* __restrict__ e) {
  __shared__ double2 v[kEvdMaxN], w[kEvdMaxN], taus[kEvdMaxN];
  __shared__ double2 part[kEvdWaves][kEvdMaxN];
  __shared__ double red[2][kEvdWaves][2];
  const int tid = static_cast<int>(threadIdx.x), lane = tid & (kWave - 1), wv = tid / kWave;
  int rb = 0;
  // Workgroup sum of (x, y).  Two alternating buffers: a buffer is written
  // again only after the next sum's barrier, which every reader has passed.
  auto block_sum = [&](double x, double y) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      x += __shfl_xor(x, off, kWave);
      y += __shfl_xor(y, off, kWave);
    }
    if (lane == 0) {
      red[rb][wv][0] = x;
      red[rb][wv][1] = y;
    }
    __syncthreads();
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int k = 0; k < kEvdWaves; ++k) {
      sx += red[rb][k][0];
      sy += red[rb][k][1];
    }
    rb ^= 1;
    return make_double2(sx, sy);
  };
  const size_t nn = static_cast<size_t>(n);
  // Lower triangle from the upper one; real diagonal.
  for (int idx = tid; idx < n * n; idx += kEvdThreads) {
    const int r = idx % n, c = idx / n;
    if (r > c) {
      const double2 u = a[c + r * nn];
      a[idx] = make_double2(u.x, -u.y);
    } else if (r == c) {
      a[idx].y = 0.0;
    }
  }
  __syncthreads();

  // ---- reduction to tridiagonal form
  for (int i = 0; i + 1 < n; ++i) {
    const int m = n - 1 - i;  // length of the column below the diagonal
    double2* colm = a + static_cast<size_t>(i) * nn + i + 1;
    double2 x = make_double2(0.0, 0.0);
    double nrm = 0.0;
    if (tid < m) {
      x = colm[tid];
      if (tid > 0) nrm = x.x * x.x + x.y * x.y;
    }
    const double2 alpha = colm[0];
    const double xnorm2 = block_sum(nrm, 0.0).x;
    double2 tau = make_double2(0.0, 0.0), scale = make_double2(0.0, 0.0);
    double beta = alpha.x;
    if (xnorm2 != 0.0 || alpha.y != 0.0) {
      const double r = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
      beta = alpha.x >= 0.0 ? -r : r;
      tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      const double2 den = make_double2(alpha.x - beta, alpha.y);  // scale = 1 / (alpha - beta)
      const double dd = den.x * den.x + den.y * den.y;
      scale = make_double2(den.x / dd, -den.y / dd);
    }
    const bool reflect = tau.x != 0.0 || tau.y != 0.0;
    if (tid < m) {
      const double2 vj = tid == 0 ? make_double2(1.0, 0.0) : (reflect ? zmul(x, scale) : x);
      v[tid] = vj;
      if (tid > 0) colm[tid] = vj;
    }
    if (tid == 0) {
      e[i] = beta;
      d[i] = a[static_cast<size_t>(i) * nn + i].x;
      taus[i] = tau;
    }
    __syncthreads();
    if (reflect) {
      // p = tau A22 v (A22 = rows and columns i + 1 .., both triangles kept).
      double2 acc[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = make_double2(0.0, 0.0);
      for (int c = wv; c < m; c += kEvdWaves) {
        const double2 vc = v[c];
        const double2* colc = a + static_cast<size_t>(i + 1 + c) * nn + i + 1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = lane + kWave * q;
          if (r < m) acc[q] = zadd(acc[q], zmul(colc[r], vc));
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = lane + kWave * q;
        if (r < m) part[wv][r] = acc[q];
      }
      __syncthreads();
      double2 pj = make_double2(0.0, 0.0), vj = make_double2(0.0, 0.0);
      if (tid < m) {
        double2 sum = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < kEvdWaves; ++k) sum = zadd(sum, part[k][tid]);
        pj = zmul(tau, sum);
        vj = v[tid];
      }
      // w = p - (tau / 2) (p^H v) v
      const double2 pv = zmulc(pj, vj);
      const double2 dot = block_sum(pv.x, pv.y);
      const double2 alpha2 = zmul(make_double2(-0.5 * tau.x, -0.5 * tau.y), dot);
      if (tid < m) w[tid] = zadd(pj, zmul(alpha2, vj));
      __syncthreads();
      // A22 -= v w^H + w v^H
      double2 vr[4], wr[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = lane + kWave * q;
        vr[q] = r < m ? v[r] : make_double2(0.0, 0.0);
        wr[q] = r < m ? w[r] : make_double2(0.0, 0.0);
      }
      for (int c = wv; c < m; c += kEvdWaves) {
        const double2 vc = v[c], wc = w[c];
        double2* colc = a + static_cast<size_t>(i + 1 + c) * nn + i + 1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = lane + kWave * q;
          if (r < m) {
            // v_r conj(w_c) + w_r conj(v_c)
            const double2 t = zadd(zmulc(wc, vr[q]), zmulc(vc, wr[q]));
            colc[r] = zsub(colc[r], t);
          }
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) d[n - 1] = a[static_cast<size_t>(n - 1) * nn + n - 1].x;
  __syncthreads();

  // ---- Q = H_0 ... H_{n-2} in place, last reflector first: column i + 1 of
  // Q is H_i e_{i+1}; columns i + 2 .. (already formed, zero in row i + 1)
  // take H_i from the left.
  for (int i = n - 2; i >= 0; --i) {
    const int m = n - 1 - i;
    const double2 tau = taus[i];
    const double2* colm = a + static_cast<size_t>(i) * nn + i + 1;
    if (tid < m) v[tid] = tid == 0 ? make_double2(1.0, 0.0) : colm[tid];
    __syncthreads();
    if (tau.x != 0.0 || tau.y != 0.0) {
      double2 ur[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = lane + kWave * q;
        ur[q] = r < m ? v[r] : make_double2(0.0, 0.0);
      }
      for (int c = i + 2 + wv; c < n; c += kEvdWaves) {
        double2* colc = a + static_cast<size_t>(c) * nn + i + 1;
        double2 t[4];
        double sx = 0.0, sy = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = lane + kWave * q;
          t[q] = r < m ? colc[r] : make_double2(0.0, 0.0);
          const double2 pr = zmulc(ur[q], t[q]);
          sx += pr.x;
          sy += pr.y;
        }
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
          sx += __shfl_xor(sx, off, kWave);
          sy += __shfl_xor(sy, off, kWave);
        }
        const double2 ts = zmul(tau, make_double2(sx, sy));
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = lane + kWave * q;
          if (r < m) colc[r] = zsub(t[q], zmul(ur[q], ts));
        }
      }
    }
    double2* colq = a + static_cast<size_t>(i + 1) * nn;
    for (int r = tid; r < n; r += kEvdThreads) {
      double2 val = make_double2(0.0, 0.0);
      if (r == i + 1) val = make_double2(1.0 - tau.x, -tau.y);
      else if (r > i + 1) val = zmul(make_double2(-tau.x, -tau.y), v[r - i - 1]);
      colq[r] = val;
    }
    __syncthreads();
  }
  for (int r = tid; r < n; r += kEvdThreads) a[r] = make_double2(r == 0 ? 1.0 : 0.0, 0.0);
}

struct SolverState {
  std::mutex lock;
  rocblas_handle handle[64] = {};
};
static SolverState g_solver;

}  // namespace rcb

using namespace rcb;

extern "C" size_t rcabeam_gram_workspace_bytes(size_t rows, size_t frames) {
  if (!rows || !frames || 2 * frames > 32 * kGramMaxColumnBlocks) return 0;
  const int nb = column_blocks(frames);
  const long long groups = (static_cast<long long>(rows) + kGramRowsPerGroup - 1) / kGramRowsPerGroup;
  const long long per_pass = triangle_blocks(nb) < kGramBlocksPerPass ? triangle_blocks(nb) : kGramBlocksPerPass;
  return static_cast<size_t>(groups * per_pass * 1024 * sizeof(float));
}

extern "C" size_t rcabeam_gram_real_elements(size_t frames) {
  const size_t ld = 32 * static_cast<size_t>(column_blocks(frames));
  return ld * ld;
}

extern "C" rcabeam_status rcabeam_iq_absmax_device(const void* iq, int storage, size_t n, unsigned* out,
                                                   int accumulate) {
  if (!iq || !out || !n || storage < 0 || storage > 2) return RCABEAM_ERROR_ARGUMENT;
  if (!accumulate) k_zero2<<<1, 64>>>(out);
  const long long blocks = static_cast<long long>((n + 255) / 256) < 4096 ? static_cast<long long>((n + 255) / 256) : 4096;
  const unsigned grid = static_cast<unsigned>(blocks);
  switch (storage) {
    case RCABEAM_STORAGE_C64:
      k_absmax<float2><<<grid, 256>>>(static_cast<const float2*>(iq), static_cast<long long>(n), out);
      break;
    case RCABEAM_STORAGE_FP16:
      k_absmax<__half2><<<grid, 256>>>(static_cast<const __half2*>(iq), static_cast<long long>(n), out);
      break;
    default:
      k_absmax<bf16x2><<<grid, 256>>>(static_cast<const bf16x2*>(iq), static_cast<long long>(n), out);
  }
  k_inverse_scale<<<1, 64>>>(out);
  return launch_status();
}

extern "C" rcabeam_status rcabeam_temporal_gram_sampled_device(const void* iq, int storage, size_t rows,
                                                               size_t row_step, size_t frames, const float* scale,
                                                               double* real_gram, float* workspace,
                                                               size_t workspace_bytes, int accumulate) {
  if (!iq || !real_gram || !workspace || !rows || !row_step || !frames || storage < 0 || storage > 2)
    return RCABEAM_ERROR_ARGUMENT;
  if (storage == RCABEAM_STORAGE_C64 && !scale) return RCABEAM_ERROR_ARGUMENT;
  if (rows > (1ull << 40) || row_step > (1ull << 20)) return RCABEAM_ERROR_ARGUMENT;
  const size_t need = rcabeam_gram_workspace_bytes(rows, frames);
  if (!need || workspace_bytes < need) return RCABEAM_ERROR_ARGUMENT;
  hipError_t e;
  const long long n = static_cast<long long>(rows), step = static_cast<long long>(row_step);
  const int t = static_cast<int>(frames);
  switch (storage) {
    case RCABEAM_STORAGE_C64:
      e = gram_launch<kSplitBf16, float2>(iq, n, step, t, scale, real_gram, workspace, accumulate);
      break;
    case RCABEAM_STORAGE_FP16:
      e = gram_launch<kExactF16, __half2>(iq, n, step, t, scale, real_gram, workspace, accumulate);
      break;
    default:
      e = gram_launch<kExactBf16, bf16x2>(iq, n, step, t, scale, real_gram, workspace, accumulate);
  }
  if (e == hipErrorInvalidValue) return RCABEAM_ERROR_ARGUMENT;
  return e == hipSuccess ? RCABEAM_SUCCESS : RCABEAM_ERROR_CUDA;
}

extern "C" rcabeam_status rcabeam_temporal_gram_device(const void* iq, int storage, size_t rows, size_t frames,
                                                       const float* scale, double* real_gram, float* workspace,
                                                       size_t workspace_bytes, int accumulate) {
  return rcabeam_temporal_gram_sampled_device(iq, storage, rows, 1, frames, scale, real_gram, workspace,
                                              workspace_bytes, accumulate);
}

extern "C" rcabeam_status rcabeam_gram_hermitian_device(const double* real_gram, size_t frames, double scale2,
                                                        void* gram) {
  if (!real_gram || !gram || !frames) return RCABEAM_ERROR_ARGUMENT;
  const long long n = static_cast<long long>(frames) * static_cast<long long>(frames);
  k_gram_form<<<dim3(static_cast<unsigned>((n + 255) / 256)), 256>>>(real_gram, column_blocks(frames),
                                                                     static_cast<int>(frames), scale2,
                                                                     static_cast<double2*>(gram));
  return launch_status();
}

namespace rcb {
// The rocBLAS / rocSOLVER handle of the current device, created on first use
// and bound to the null stream (nullptr on failure).
rocblas_handle solver_handle() {
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= 64) return nullptr;
  std::lock_guard<std::mutex> guard(g_solver.lock);
  if (!g_solver.handle[device]) {
    if (rocblas_create_handle(&g_solver.handle[device]) != rocblas_status_success) return nullptr;
    rocblas_set_stream(g_solver.handle[device], nullptr);
  }
  return g_solver.handle[device];
}
}  // namespace rcb

extern "C" rcabeam_status rcabeam_eigh_device(void* gram, double* values, double* workspace_e, int* info,
                                              size_t frames) {
  if (!gram || !values || !workspace_e || !info || !frames || frames > 0x7fffffff) return RCABEAM_ERROR_ARGUMENT;
  rocblas_handle handle = solver_handle();
  if (!handle) return RCABEAM_ERROR_CUDA;
  const rocblas_int n = static_cast<rocblas_int>(frames);
  rocblas_status s;
  if (frames <= static_cast<size_t>(kEvdMaxN)) {
    k_hetrd_q<<<1, kEvdThreads>>>(static_cast<double2*>(gram), static_cast<int>(frames), values, workspace_e);
    if (hipGetLastError() != hipSuccess) return RCABEAM_ERROR_CUDA;
    s = rocsolver_zstedc(handle, rocblas_evect_original, n, values, workspace_e,
                         static_cast<rocblas_double_complex*>(gram), n, info);
  } else {
    const long long count = static_cast<long long>(frames) * static_cast<long long>(frames);
    if ((count + 255) / 256 > 0xffffffffll) return RCABEAM_ERROR_ARGUMENT;
    k_hermitian_from_upper<<<static_cast<unsigned>((count + 255) / 256), 256>>>(static_cast<double2*>(gram), n);
    if (hipGetLastError() != hipSuccess) return RCABEAM_ERROR_CUDA;
    s = rocsolver_zheevd(handle, rocblas_evect_original, rocblas_fill_upper, n,
                         static_cast<rocblas_double_complex*>(gram), n, values, workspace_e, info);
  }
  if (s == rocblas_status_invalid_size || s == rocblas_status_invalid_pointer) return RCABEAM_ERROR_ARGUMENT;
  return s == rocblas_status_success ? RCABEAM_SUCCESS : RCABEAM_ERROR_CUDA;
}

template <bool FILTER>
static rcabeam_status project(const void* iq, int storage, size_t rows, size_t frames, const void* vectors, size_t r,
                              int subtract, double* power, void* filtered) {
  if (!iq || !rows || !frames || frames > 1024 || (r && !vectors) || storage < 0 || storage > 2)
    return RCABEAM_ERROR_ARGUMENT;
  const unsigned grid = static_cast<unsigned>((rows + 3) / 4);
  const long long n = static_cast<long long>(rows);
  const int t = static_cast<int>(frames);
  const double2* v = static_cast<const double2*>(vectors);
  float2* out = static_cast<float2*>(filtered);
  const int ri = static_cast<int>(r);
  if (frames <= 256 && r <= static_cast<size_t>(kProjectMaxR)) {
    // Short rows, few vectors: 8 or 16 lanes per row, vectors in LDS.
    const int lanes = frames <= 128 ? 8 : 16;
    const unsigned blocks = static_cast<unsigned>((rows + 256 / lanes - 1) / (256 / lanes));
    const size_t lds = std::max<size_t>(r * frames * sizeof(double2), sizeof(double2));
#define RCB_PROJECT8(TYPE, L, PER)                                                                           \
  do {                                                                                                        \
    if (ri <= 4)                                                                                              \
      k_project_rows8<TYPE, FILTER, L, PER, 4><<<blocks, 256, lds>>>(static_cast<const TYPE*>(iq), n, t, v, ri, \
                                                                     subtract, power, out);                  \
    else                                                                                                      \
      k_project_rows8<TYPE, FILTER, L, PER, kProjectMaxR><<<blocks, 256, lds>>>(static_cast<const TYPE*>(iq), n, \
                                                                                t, v, ri, subtract, power, out); \
  } while (0)
#define RCB_PROJECT8_SHAPES(TYPE)                    \
  if (frames <= 64) RCB_PROJECT8(TYPE, 8, 8);         \
  else if (frames <= 128) RCB_PROJECT8(TYPE, 8, 16);  \
  else if (frames <= 208) RCB_PROJECT8(TYPE, 16, 13); \
  else RCB_PROJECT8(TYPE, 16, 16)
    switch (storage) {
      case RCABEAM_STORAGE_C64:
        RCB_PROJECT8_SHAPES(float2);
        break;
      case RCABEAM_STORAGE_FP16:
        RCB_PROJECT8_SHAPES(__half2);
        break;
      default:
        RCB_PROJECT8_SHAPES(bf16x2);
    }
#undef RCB_PROJECT8_SHAPES
#undef RCB_PROJECT8
    return launch_status();
  }
#define RCB_PROJECT(TYPE, MPL) \
  k_project<TYPE, FILTER, MPL><<<grid, 256>>>(static_cast<const TYPE*>(iq), n, t, v, ri, subtract, power, out)
  const bool small = frames <= 4 * kWave;
  switch (storage) {
    case RCABEAM_STORAGE_C64:
      if (small) RCB_PROJECT(float2, 4); else RCB_PROJECT(float2, 16);
      break;
    case RCABEAM_STORAGE_FP16:
      if (small) RCB_PROJECT(__half2, 4); else RCB_PROJECT(__half2, 16);
      break;
    default:
      if (small) RCB_PROJECT(bf16x2, 4); else RCB_PROJECT(bf16x2, 16);
  }
#undef RCB_PROJECT
  return launch_status();
}

extern "C" rcabeam_status rcabeam_filtered_power_device(const void* iq, int storage, size_t rows, size_t frames,
                                                        const void* vectors, size_t r, int subtract, double* power) {
  if (!power) return RCABEAM_ERROR_ARGUMENT;
  return project<false>(iq, storage, rows, frames, vectors, r, subtract, power, nullptr);
}

extern "C" rcabeam_status rcabeam_temporal_filter_device(const void* iq, int storage, size_t rows, size_t frames,
                                                         const void* vectors, size_t r, int subtract, void* out) {
  if (!out) return RCABEAM_ERROR_ARGUMENT;
  return project<true>(iq, storage, rows, frames, vectors, r, subtract, nullptr, out);
}

// ---- exact Gram without allocation -----------------------------------------
// The partial sums live in this file-scope device array instead of a caller
// workspace (each GPU the library runs on holds its own copy until the process
// ends).  Host code never passes the array's name to a kernel: its device
// address on the current GPU comes from hipGetSymbolAddress, looked up once per
// device on the entry point's first call there.  Every partial a call reads is
// written earlier in the same call, so nothing needs clearing; calls that
// overlap on one GPU share the array and race.
constexpr size_t kGramScratchBytes = RCABEAM_GRAM_F64_SCRATCH_BYTES;
static __device__ __align__(256) double scratch[kGramScratchBytes / sizeof(double)];

struct ScratchState {
  std::mutex lock;
  double* ptr[64] = {};
};
static ScratchState g_scratch;

static hipError_t scratch_pointer(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> guard(g_scratch.lock);
  if (!g_scratch.ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch.ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch.ptr[dev];
  return hipSuccess;
}

extern "C" size_t rcabeam_gram_f64_workspace_bytes(size_t rows, size_t frames) {
  if (!rows || !gram64_frames_ok(frames)) return 0;
  return gram64_partial_bytes(gram64_shape(rows, frames, kG64RowsPerGroup));
}

extern "C" rcabeam_status rcabeam_temporal_gram_f64_device(const void* iq, size_t rows, size_t frames,
                                                           double* real_gram, double* workspace,
                                                           size_t workspace_bytes, int accumulate) {
  if (!iq || !real_gram || !workspace || !rows || !frames) return RCABEAM_ERROR_ARGUMENT;
  const size_t need = rcabeam_gram_f64_workspace_bytes(rows, frames);
  if (!need || workspace_bytes < need) return RCABEAM_ERROR_ARGUMENT;
  const Gram64Shape sh = gram64_shape(rows, frames, kG64RowsPerGroup);
  return gram64_launch(static_cast<const float2*>(iq), rows, frames, sh, workspace, real_gram, accumulate) == hipSuccess
             ? RCABEAM_SUCCESS
             : RCABEAM_ERROR_CUDA;
}

extern "C" rcabeam_status rcabeam_temporal_gram_f64_noalloc_device(const void* iq, size_t rows, size_t frames,
                                                                   double* real_gram, int accumulate) {
  if (!iq || !real_gram || !rows || !gram64_frames_ok(frames)) return RCABEAM_ERROR_ARGUMENT;
  Gram64Shape sh = gram64_shape(rows, frames, kG64RowsPerGroup);
  const size_t per_group = static_cast<size_t>(sh.total) * 256 * sizeof(double);
  const long long max_groups = static_cast<long long>(kGramScratchBytes / per_group);
  if (max_groups < 1) return RCABEAM_ERROR_ARGUMENT;
  if (sh.groups > max_groups) {
    // Larger row groups so that the partials fit (float64 summation order then differs from the
    // workspace entry point's by regrouping only).
    long long per = (static_cast<long long>(rows) + max_groups - 1) / max_groups;
    per = (per + kG64MaxRows - 1) / kG64MaxRows * kG64MaxRows;
    sh = gram64_shape(rows, frames, per);
  }
  double* partials = nullptr;
  if (scratch_pointer(&partials) != hipSuccess) return RCABEAM_ERROR_CUDA;
  return gram64_launch(static_cast<const float2*>(iq), rows, frames, sh, partials, real_gram, accumulate) == hipSuccess
             ? RCABEAM_SUCCESS
             : RCABEAM_ERROR_CUDA;
}
