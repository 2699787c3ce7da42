// Direct temporal SVD and complex128 clutter filtering of a Casorati matrix on
// MI355X: the direct-SVD path of reference src/rcabeam/gpu_clutter.py:43-109
// (cuSOLVER gesvd on a column-major copy) and :276-319 (U Sigma V^H
// reconstruction of the retained interval), and the complex128 Gram path of
// the NumPy reference src/rcabeam/clutter.py:18-44, 97-101 for complex128
// input.  Dense linear algebra runs in rocSOLVER (gesvd, on the handle that
// rcabeam_eigh_device creates) and rocBLAS (herk, gemm); the kernels here move
// data between the row-major (rows, frames) layout of the IQ and the
// column-major layout the solvers take, through 32 x 32 LDS tiles so that
// both sides of each copy are coalesced.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <mutex>

#include "common.hpp"
#include "rcabeam_mi355x.h"

namespace rcb {

rocblas_handle solver_handle();  // clutter.hip: one per device, on the null stream

namespace {

constexpr int kTile = 32;

template <typename T>
__device__ __forceinline__ double2 widen2(T v);
template <>
__device__ __forceinline__ double2 widen2<float2>(float2 v) {
  return make_double2(v.x, v.y);
}
template <>
__device__ __forceinline__ double2 widen2<double2>(double2 v) {
  return v;
}

template <typename T>
__device__ __forceinline__ T narrow2(double2 v);
template <>
__device__ __forceinline__ float2 narrow2<float2>(double2 v) {
  return make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
}
template <>
__device__ __forceinline__ double2 narrow2<double2>(double2 v) {
  return v;
}

// Row-major x (rows, frames) -> column-major m rows [offset, offset + rows)
// (leading dimension ld).  One 32 x 32 tile per 256-thread block.
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_to_colmajor(const TI* __restrict__ x, long long rows, int frames,
                                                     TO* __restrict__ m, long long ld, long long offset) {
  __shared__ double2 tile[kTile][kTile + 1];
  const long long r0 = static_cast<long long>(blockIdx.x) * kTile;
  const int t0 = static_cast<int>(blockIdx.y) * kTile;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < kTile; k += 8) {
    const long long r = r0 + k;
    const int t = t0 + tx;
    tile[k][tx] = (r < rows && t < frames) ? widen2(x[r * frames + t]) : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int k = ty; k < kTile; k += 8) {
    const int t = t0 + k;
    const long long r = r0 + tx;
    if (r < rows && t < frames) m[static_cast<long long>(t) * ld + offset + r] = narrow2<TO>(tile[tx][k]);
  }
}

// U[:, k] *= sigma_k for k in [start, stop), 0 otherwise (rows [0, rows) of a
// column-major matrix with leading dimension ld).
template <typename T, typename R>
__global__ __launch_bounds__(256) void k_scale_columns(T* __restrict__ u, long long rows, long long ld, int frames,
                                                       const R* __restrict__ values, int start, int stop) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows * frames) return;
  const int k = static_cast<int>(i / rows);
  const long long r = i - static_cast<long long>(k) * rows;
  T* p = u + static_cast<long long>(k) * ld + r;
  const R s = (k >= start && k < stop) ? values[k] : R(0);
  if (s == R(0)) {
    p->x = R(0);  // columns past the solver's min(rows, frames) are not singular vectors
    p->y = R(0);
  } else {
    p->x *= s;
    p->y *= s;
  }
}

// Copy a row-major complex block to the output type.
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_convert(const TI* __restrict__ in, long long n, TO* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = narrow2<TO>(widen2(in[i]));
}

__global__ __launch_bounds__(256) void k_zero(double2* __restrict__ out, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double2(0.0, 0.0);
}

template <typename TI, typename TO>
hipError_t to_colmajor(const void* iq, size_t rows, size_t frames, void* matrix, size_t ld, size_t offset) {
  const dim3 grid(static_cast<unsigned>((rows + kTile - 1) / kTile), static_cast<unsigned>((frames + kTile - 1) / kTile));
  k_to_colmajor<TI, TO><<<grid, 256>>>(static_cast<const TI*>(iq), static_cast<long long>(rows),
                                       static_cast<int>(frames), static_cast<TO*>(matrix), static_cast<long long>(ld),
                                       static_cast<long long>(offset));
  return hipGetLastError();
}

bool fits_int(size_t v) { return v <= 0x7fffffff; }

}  // namespace
}  // namespace rcb

using namespace rcb;

extern "C" rcabeam_status rcabeam_casorati_colmajor_device(const void* iq, int complex128_in, size_t rows,
                                                           size_t frames, void* matrix, int complex128_out, size_t ld,
                                                           size_t offset) {
  if (!iq || !matrix || !rows || !frames || !fits_int(frames) || ld < offset + rows || rows > (1ull << 40))
    return RCABEAM_ERROR_ARGUMENT;
  if ((rows + kTile - 1) / kTile > 0x7fffffff) return RCABEAM_ERROR_ARGUMENT;
  hipError_t e;
  if (complex128_in)
    e = complex128_out ? to_colmajor<double2, double2>(iq, rows, frames, matrix, ld, offset)
                       : to_colmajor<double2, float2>(iq, rows, frames, matrix, ld, offset);
  else
    e = complex128_out ? to_colmajor<float2, double2>(iq, rows, frames, matrix, ld, offset)
                       : to_colmajor<float2, float2>(iq, rows, frames, matrix, ld, offset);
  return e == hipSuccess ? RCABEAM_SUCCESS : RCABEAM_ERROR_CUDA;
}

extern "C" rcabeam_status rcabeam_svd_device(void* matrix, int complex128, size_t rows, size_t frames, int want_u,
                                             void* values, void* vh, void* e_work, int* info) {
  if (!matrix || !values || !vh || !e_work || !info || !rows || !frames || !fits_int(rows) || !fits_int(frames))
    return RCABEAM_ERROR_ARGUMENT;
  rocblas_handle handle = solver_handle();
  if (!handle) return RCABEAM_ERROR_CUDA;
  const rocblas_int m = static_cast<rocblas_int>(rows), n = static_cast<rocblas_int>(frames);
  const rocblas_svect left = want_u ? rocblas_svect_overwrite : rocblas_svect_none;
  rocblas_status s;
  if (complex128)
    s = rocsolver_zgesvd(handle, left, rocblas_svect_all, m, n, static_cast<rocblas_double_complex*>(matrix), m,
                         static_cast<double*>(values), nullptr, m, static_cast<rocblas_double_complex*>(vh), n,
                         static_cast<double*>(e_work), rocblas_outofplace, info);
  else
    s = rocsolver_cgesvd(handle, left, rocblas_svect_all, m, n, static_cast<rocblas_float_complex*>(matrix), m,
                         static_cast<float*>(values), nullptr, m, static_cast<rocblas_float_complex*>(vh), n,
                         static_cast<float*>(e_work), rocblas_outofplace, info);
  if (s == rocblas_status_invalid_size || s == rocblas_status_invalid_pointer) return RCABEAM_ERROR_ARGUMENT;
  return s == rocblas_status_success ? RCABEAM_SUCCESS : RCABEAM_ERROR_CUDA;
}

extern "C" rcabeam_status rcabeam_scale_columns_device(void* u, int complex128, size_t rows, size_t ld, size_t frames,
                                                       const void* values, size_t start, size_t stop) {
  if (!u || !values || !rows || !frames || ld < rows || start > stop || stop > frames || !fits_int(frames))
    return RCABEAM_ERROR_ARGUMENT;
  const long long n = static_cast<long long>(rows) * static_cast<long long>(frames);
  const long long blocks = (n + 255) / 256;
  if (blocks > 0x7fffffffLL) return RCABEAM_ERROR_ARGUMENT;
  if (complex128)
    k_scale_columns<double2, double><<<static_cast<unsigned>(blocks), 256>>>(
        static_cast<double2*>(u), static_cast<long long>(rows), static_cast<long long>(ld), static_cast<int>(frames),
        static_cast<const double*>(values), static_cast<int>(start), static_cast<int>(stop));
  else
    k_scale_columns<float2, float><<<static_cast<unsigned>(blocks), 256>>>(
        static_cast<float2*>(u), static_cast<long long>(rows), static_cast<long long>(ld), static_cast<int>(frames),
        static_cast<const float*>(values), static_cast<int>(start), static_cast<int>(stop));
  return launch_status();
}

extern "C" rcabeam_status rcabeam_times_square_device(const void* b, int complex128, size_t rows, size_t ld,
                                                      size_t frames, const void* a, void* out, int out_complex128,
                                                      void* workspace, size_t chunk_rows) {
  if (!b || !a || !out || !rows || !frames || ld < rows || !fits_int(ld) || !fits_int(frames) || !chunk_rows ||
      !fits_int(chunk_rows))
    return RCABEAM_ERROR_ARGUMENT;
  const bool direct = out_complex128 == complex128;  // the product lands in out in its own type
  if (!direct && !workspace) return RCABEAM_ERROR_ARGUMENT;
  rocblas_handle handle = solver_handle();
  if (!handle) return RCABEAM_ERROR_CUDA;
  const rocblas_int t = static_cast<rocblas_int>(frames);
  // out (rows, frames) row-major = (B A) row-major = column-major (frames, rows) matrix A^T B^T.
  for (size_t r0 = 0; r0 < rows; r0 += chunk_rows) {
    const size_t n = rows - r0 < chunk_rows ? rows - r0 : chunk_rows;
    const rocblas_int nn = static_cast<rocblas_int>(n);
    rocblas_status s;
    if (complex128) {
      const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
      auto* c = direct ? static_cast<rocblas_double_complex*>(out) + r0 * frames
                       : static_cast<rocblas_double_complex*>(workspace);
      s = rocblas_zgemm(handle, rocblas_operation_transpose, rocblas_operation_transpose, t, nn, t, &one,
                        static_cast<const rocblas_double_complex*>(a), t,
                        static_cast<const rocblas_double_complex*>(b) + r0, static_cast<rocblas_int>(ld), &zero, c, t);
    } else {
      const rocblas_float_complex one{1.f, 0.f}, zero{0.f, 0.f};
      auto* c = direct ? static_cast<rocblas_float_complex*>(out) + r0 * frames
                       : static_cast<rocblas_float_complex*>(workspace);
      s = rocblas_cgemm(handle, rocblas_operation_transpose, rocblas_operation_transpose, t, nn, t, &one,
                        static_cast<const rocblas_float_complex*>(a), t,
                        static_cast<const rocblas_float_complex*>(b) + r0, static_cast<rocblas_int>(ld), &zero, c, t);
    }
    if (s != rocblas_status_success) return RCABEAM_ERROR_CUDA;
    if (!direct) {
      const long long count = static_cast<long long>(n) * static_cast<long long>(frames);
      const unsigned blocks = static_cast<unsigned>((count + 255) / 256);
      if (complex128)
        k_convert<double2, float2><<<blocks, 256>>>(static_cast<const double2*>(workspace), count,
                                                    static_cast<float2*>(out) + r0 * frames);
      else
        k_convert<float2, double2><<<blocks, 256>>>(static_cast<const float2*>(workspace), count,
                                                    static_cast<double2*>(out) + r0 * frames);
      if (hipGetLastError() != hipSuccess) return RCABEAM_ERROR_CUDA;
    }
  }
  return RCABEAM_SUCCESS;
}

extern "C" rcabeam_status rcabeam_gram_colmajor_device(const void* matrix, size_t rows, size_t ld, size_t frames,
                                                       void* gram) {
  if (!matrix || !gram || !rows || !frames || ld < rows || !fits_int(ld) || !fits_int(rows) || !fits_int(frames))
    return RCABEAM_ERROR_ARGUMENT;
  rocblas_handle handle = solver_handle();
  if (!handle) return RCABEAM_ERROR_CUDA;
  const double one = 1.0, zero = 0.0;
  const rocblas_status s = rocblas_zherk(handle, rocblas_fill_upper, rocblas_operation_conjugate_transpose,
                                         static_cast<rocblas_int>(frames), static_cast<rocblas_int>(rows), &one,
                                         static_cast<const rocblas_double_complex*>(matrix),
                                         static_cast<rocblas_int>(ld), &zero,
                                         static_cast<rocblas_double_complex*>(gram), static_cast<rocblas_int>(frames));
  return s == rocblas_status_success ? RCABEAM_SUCCESS : RCABEAM_ERROR_CUDA;
}

extern "C" rcabeam_status rcabeam_projector_device(const void* basis, size_t frames, size_t start, size_t stop,
                                                   void* projector) {
  if (!basis || !projector || !frames || start > stop || stop > frames || !fits_int(frames))
    return RCABEAM_ERROR_ARGUMENT;
  rocblas_handle handle = solver_handle();
  if (!handle) return RCABEAM_ERROR_CUDA;
  const rocblas_int t = static_cast<rocblas_int>(frames), r = static_cast<rocblas_int>(stop - start);
  if (r == 0) {
    const long long n = static_cast<long long>(frames) * static_cast<long long>(frames);
    k_zero<<<static_cast<unsigned>((n + 255) / 256), 256>>>(static_cast<double2*>(projector), n);
    return launch_status();
  }
  const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
  const auto* v = static_cast<const rocblas_double_complex*>(basis) + start * frames;
  const rocblas_status s = rocblas_zgemm(handle, rocblas_operation_none, rocblas_operation_conjugate_transpose, t, t, r,
                                         &one, v, t, v, t, &zero, static_cast<rocblas_double_complex*>(projector), t);
  return s == rocblas_status_success ? RCABEAM_SUCCESS : RCABEAM_ERROR_CUDA;
}
