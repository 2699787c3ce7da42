/* MI355X extensions of the rcabeam C ABI.
 *
 * These entry points have no counterpart in the reference's C header; each
 * comment names the reference code whose job it takes over.  Conventions match
 * include/rcabeam.h: device pointers, default (null) stream, asynchronous
 * return after enqueue, launch/argument errors only, no allocation unless a
 * comment says otherwise.
 */
#pragma once

#include <stddef.h>

#include "rcabeam.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Sample storage of packed channels or beamformed IQ. */
typedef enum rcabeam_storage {
    RCABEAM_STORAGE_C64 = 0,  /* complex64: two float32 */
    RCABEAM_STORAGE_FP16 = 1, /* two IEEE binary16 (real, imaginary) */
    RCABEAM_STORAGE_BF16 = 2, /* two bfloat16, real in the low half-word */
    RCABEAM_STORAGE_FP16_SPLIT = 3, /* packed channels only: complex64 as two FP16 pairs (hi, lo), 8 bytes,
                                       see rcabeam_pack_ensemble_split_device */
} rcabeam_storage;

/* Pack complex64 channels to complex BF16, same layout conversion and range
 * flag semantics as rcabeam_pack_ensemble_fp16_device (non-finite components
 * or components beyond the largest finite bfloat16 set *range_error = 1).
 * Extends reference src/rcabeam/ensemble.cu:129-141 to bfloat16. */
rcabeam_status rcabeam_pack_ensemble_bf16_device(const void* source, void* target, int* range_error, size_t ns,
                                                 size_t nc, size_t na, size_t nt);

/* OPW from packed BF16 RC/CR, float32 arithmetic; arguments as
 * rcabeam_opw_fp16_device without the unused geometry workspace. */
rcabeam_status rcabeam_opw_bf16_device(const void* rc, const void* cr, const float* scan, const float* xe,
                                       const float* ye, const float* angles, const float* starts, float* pd,
                                       void* signal, size_t ns, size_t nc, size_t na, size_t nt, size_t nv, float c,
                                       float fs, float fd, float fn);

/* Tap-coefficient precision of the 16-bit delay-and-sum. */
typedef enum rcabeam_coefficients {
    RCABEAM_COEF_FLOAT32 = 0, /* float32-accurate: samples widened to float32, or FP16 samples times
                                 hi + lo FP16 coefficients (exact products, float32 accumulation) */
    RCABEAM_COEF_16BIT = 1,   /* one coefficient of the channel storage type per sample
                                 (interpolation weight x phasor), exact 16-bit products,
                                 float32 accumulation: half the arithmetic of RCABEAM_COEF_FLOAT32 */
} rcabeam_coefficients;

/* OPW with any channel storage and any beamformed-IQ storage (rcabeam_storage
 * codes); float32 accumulation, coefficient precision per rcabeam_coefficients
 * (ignored for complex64 channels).  signal (optional) holds (nv, nt) samples
 * of signal_storage; pd as rcabeam_ensemble_device.  The 16-bit signal is the
 * mixed-precision form of the clutter-filter input.  Extends reference
 * src/rcabeam/ensemble.cu:798-809 (rcabeam_opw_fp16_device). */
rcabeam_status rcabeam_opw_mixed_device(const void* rc, const void* cr, const float* scan, const float* xe,
                                        const float* ye, const float* angles, const float* starts, float* pd,
                                        void* signal, int channel_storage, int signal_storage, int coefficients,
                                        size_t ns, size_t nc, size_t na, size_t nt, size_t nv, float c, float fs,
                                        float fd, float fn);

/* OPW (method 0) or separate-aperture IQ (method 5, complex64 channels, signal
 * (nv, 2, nt) complex64 RC then CR) on a Cartesian grid: axes x (nx), z (nz),
 * y (ny); voxel id (ix * nz + iz) * ny + iy as _scan_coords orders them, so
 * the outputs are those of rcabeam_opw_mixed_device / rcabeam_ensemble_device
 * on the scan list of the grid.  Tiles are 8 x 8 lateral patches at one depth
 * (short sample windows; the fastest form).  pd: 2 * nx*nz*ny floats; signal
 * (optional for method 0): (nv, nt) samples of signal_storage.  Requires
 * na <= 16 and nt <= 208 (RCABEAM_ERROR_ARGUMENT otherwise).  Other arguments
 * as rcabeam_opw_mixed_device. */
rcabeam_status rcabeam_opw_grid_device(const void* rc, const void* cr, int channel_storage, int coefficients,
                                       const float* x, const float* z, const float* y, size_t nx, size_t nz,
                                       size_t ny, const float* xe, const float* ye, const float* angles,
                                       const float* starts, float* pd, void* signal, int signal_storage, int method,
                                       size_t ns, size_t nc, size_t na, size_t nt, float c, float fs, float fd,
                                       float fn);

/* Float32-accurate channels for the MFMA delay-and-sum: complex64 (samples,
 * channels, angles, frames) packed to (channels, angles, samples, frames) as
 * word pairs (hi, lo) of FP16 complex values, hi = fp16(x 2^e) and
 * lo = fp16(x 2^e - hi), 8 bytes per sample (as complex64).  One power-of-two
 * scale 2^e per ensemble: absmax points at out[0] of
 * rcabeam_iq_absmax_device over both apertures (the joint max |component| m),
 * and e puts m 2^e in [2^14, 2^15) (e = 0 when m is 0 or not finite).
 * Components within 2^17 of m keep 22 significant bits; the rest an absolute
 * error below 2^-40 m.  No reference counterpart (it replaces complex64
 * packing, ensemble.cu:119-128, for the MFMA path). */
rcabeam_status rcabeam_pack_ensemble_split_device(const void* source, void* target, const unsigned* absmax,
                                                  size_t ns, size_t nc, size_t na, size_t nt);

/* OPW on a Cartesian grid from split FP16 channels (both apertures packed with
 * the same absmax) on MFMA: hi and lo coefficients times hi and lo samples
 * (three exact FP16 products per term, float32 accumulation; the dropped
 * lo x lo product is below 2^-22 of the term), then the scale 2^-e undone
 * exactly.  Float32-accurate: the same IQ and power as the complex64 grid
 * kernel up to float32 rounding.  Arguments as rcabeam_opw_grid_device
 * (method 0 only; nt <= 224, na <= 16, 8-byte aligned channels). */
rcabeam_status rcabeam_opw_grid_split_device(const void* rc, const void* cr, const unsigned* absmax, const float* x,
                                             const float* z, const float* y, size_t nx, size_t nz, size_t ny,
                                             const float* xe, const float* ye, const float* angles,
                                             const float* starts, float* pd, void* signal, int signal_storage,
                                             size_t ns, size_t nc, size_t na, size_t nt, float c, float fs, float fd,
                                             float fn);

/* Grid OPW (method 0, complex64 channels, the VALU grid kernel) over one
 * aperture only: aperture 0 sums the RC channels, 1 the CR channels.  With
 * accumulate != 0 the sums are added to the complex64 signal already there
 * (the other aperture's) and pd is the power of the total; otherwise signal is
 * overwritten.  RC then CR with accumulate gives rcabeam_opw_grid_device's IQ
 * and power (up to float32 rounding of the one extra addition), so the upload
 * of CR can overlap the RC pass.  signal: (nv, nt) complex64, required. */
rcabeam_status rcabeam_opw_grid_aperture_device(const void* rc, const void* cr, const float* x, const float* z,
                                                const float* y, size_t nx, size_t nz, size_t ny, const float* xe,
                                                const float* ye, const float* angles, const float* starts, float* pd,
                                                void* signal, int aperture, int accumulate, size_t ns, size_t nc,
                                                size_t na, size_t nt, float c, float fs, float fd, float fn);

/* Inverse of rcabeam_pack_ensemble_device: complex64 (channels, angles,
 * samples, frames) back to (samples, channels, angles, frames).  Used to
 * return channel gradients in the caller's layout. */
rcabeam_status rcabeam_unpack_ensemble_device(const void* source, void* target, size_t ns, size_t nc, size_t na,
                                              size_t nt);

/* Host ingest (no reference counterpart; replaces the pageable upload of
 * src/rcabeam/bindings/python.cpp:527-531): copy `bytes` from pageable host
 * memory to device memory through two 64 MiB pinned staging buffers, host
 * threads filling one buffer while the other drains to the device.  threads:
 * host copy threads (1-64).  Returns after the copy has completed.  The
 * staging buffers and a stream are created per device on first use and kept
 * until the process ends (128 MiB of pinned host memory per device); calls on
 * one device are serialized. */
rcabeam_status rcabeam_upload_host(const void* host, void* device, size_t bytes, int threads);

/* As rcabeam_upload_host, ordered after work already enqueued on the null
 * stream (which may still read or write `device`).  The _concurrent form
 * does not wait for that work, so the copy overlaps kernels already in flight
 * (e.g. the delay-and-sum of the other aperture); the caller guarantees that
 * no queued work touches `device`. */
rcabeam_status rcabeam_upload_host_concurrent(const void* host, void* device, size_t bytes, int threads);

/* XDoppler power from separate-aperture IQ: apertures (nv, 2, nt) complex64
 * (RC then CR per voxel, as rcabeam_opw_grid_device method 5 writes them);
 * pd[2v] = |mean_t RC_t conj(CR_t)|, pd[2v + 1] = its imaginary part;
 * signal (optional, (nv, nt) complex64) = RC_t conj(CR_t).  With method 5 on
 * a grid it gives rcabeam_ensemble_device method 1 (reference
 * src/rcabeam/ensemble.cu:236-315, 778-780) up to float32 summation order. */
rcabeam_status rcabeam_xdoppler_from_apertures_device(const void* apertures, size_t nv, size_t nt, float* pd,
                                                      void* signal);

/* Adjoint (transpose) of rcabeam_opw_grid_device for complex64 channels, the
 * backward pass of a differentiable OPW (no reference counterpart).  rc, cr:
 * the forward's packed complex64 channels; grad_iq: (nv, nt) complex64
 * gradient dL/dRe(IQ) + i dL/dIm(IQ) of a real loss.  Accumulates (the
 * caller zeroes them) grad_rc, grad_cr: packed (channels, angles, samples,
 * frames) complex64 gradients in the same convention, and *grad_c (float64):
 * dL/dc.  Taps are recomputed (nothing is stored between the passes), frames
 * in chunks of 32; na <= 16.  Other arguments as rcabeam_opw_grid_device. */
rcabeam_status rcabeam_opw_adjoint_grid_device(const void* rc, const void* cr, const float* x, const float* z,
                                               const float* y, size_t nx, size_t nz, size_t ny, const float* xe,
                                               const float* ye, const float* angles, const float* starts,
                                               const void* grad_iq, void* grad_rc, void* grad_cr, double* grad_c,
                                               size_t ns, size_t nc, size_t na, size_t nt, float c, float fs,
                                               float fd, float fn);

/* The same adjoint, bit-identical from run to run: each workgroup's gradient
 * sums are added as 64-bit fixed-point integers (multiples of 2^-e, e from
 * max |grad_iq| so that no sample's total can overflow; integer addition does
 * not depend on the order of the workgroups), converted to float32 and added
 * to grad_rc / grad_cr at the end, and dL/dc is summed from per-workgroup
 * partials in a fixed order.  workspace: 256-byte aligned, at least
 * rcabeam_opw_adjoint_workspace_bytes(nx, nz, ny, ns, nc, na, nt) bytes (the
 * fixed-point planes: 32 bytes per packed channel sample of both apertures). */
size_t rcabeam_opw_adjoint_workspace_bytes(size_t nx, size_t nz, size_t ny, size_t ns, size_t nc, size_t na,
                                           size_t nt);
rcabeam_status rcabeam_opw_adjoint_grid_deterministic_device(
    const void* rc, const void* cr, const float* x, const float* z, const float* y, size_t nx, size_t nz, size_t ny,
    const float* xe, const float* ye, const float* angles, const float* starts, const void* grad_iq, void* grad_rc,
    void* grad_cr, double* grad_c, void* workspace, size_t workspace_bytes, size_t ns, size_t nc, size_t na,
    size_t nt, float c, float fs, float fd, float fn);

/* Adds to *out (a device counter the caller zeroes) the number of
 * (voxel, aperture, channel, angle) taps inside the dynamic aperture and the
 * closed sample window: the work the ensemble kernels perform.  Used for
 * algorithmic-flop accounting; arguments as rcabeam_ensemble_device. */
rcabeam_status rcabeam_count_active_taps_device(const float* scan, const float* xe, const float* ye,
                                                const float* angles, const float* starts, size_t ns, size_t nc,
                                                size_t na, size_t nv, float c, float fs, float fd, float fn,
                                                unsigned long long* out);

/* ---- low-rank + residual coding of 16-bit channels (mixed precision; no
 * reference counterpart) -------------------------------------------------
 * Pack complex64 (samples, channels, angles, frames) channels to the packed
 * (channels, angles, samples, frames + 2k) layout in `storage` (FP16 or
 * BF16): with basis V (complex128, k rows of nt = the columns of an nt x k
 * matrix with orthonormal columns) each row c is written as its residual
 * e = c - (c V) V^H in frames [0, nt) and the coefficients w = c V as 16-bit
 * hi and lo parts in frames nt + 2j, nt + 2j + 1.  Delay-and-sum of the
 * coded channels (nt + 2k frames, any rcabeam_opw_grid_device form) followed
 * by rcabeam_lowrank_expand_device gives the delay-and-sum of c with only
 * the residual rounded to 16 bits.  k <= 12, nt <= 256; range_error as
 * rcabeam_pack_ensemble_fp16_device. */
rcabeam_status rcabeam_pack_ensemble_lowrank_device(const void* source, void* target, int storage,
                                                    const void* basis, size_t k, int* range_error, size_t ns,
                                                    size_t nc, size_t na, size_t nt);

/* Expand coded delay-and-sum output coded (nv, nt + 2k) complex64 into
 * x (nv, nt) of out_storage: x_t = d_t + sum_j (d_{nt+2j} + d_{nt+2j+1})
 * conj(V[t, j]); pd (optional, 2 nv floats): mean |x_t|^2 in the even
 * entries, 0 in the odd ones.  out may be null when pd is given. */
rcabeam_status rcabeam_lowrank_expand_device(const void* coded, const void* basis, size_t k, void* out,
                                             int out_storage, float* pd, size_t nv, size_t nt);

/* Orthonormal basis (complex128, k rows of `frames` = the columns of a
 * frames x k matrix) of the dominant k-dimensional invariant subspace of the
 * Hermitian frames x frames complex128 matrix `gram` (column-major, both
 * triangles), by `iterations` rounds of subspace iteration from k fixed
 * pseudo-random vectors (modified Gram-Schmidt after each product; a zero
 * matrix gives a zero basis).  One workgroup, no allocation.  The span converges at the rate (lambda_{k+1} / lambda_k) per
 * round; the low-rank channel coding uses it for its basis (any orthonormal
 * basis codes exactly; one that catches the strong modes keeps the 16-bit
 * residual small).  No reference counterpart.  1 <= k <= 12, k < frames,
 * k * frames <= 4096, 1 <= iterations <= 1000, 16-byte aligned buffers. */
rcabeam_status rcabeam_dominant_subspace_device(const void* gram, size_t frames, size_t k, size_t iterations,
                                                void* basis);

/* ---- temporal clutter filter (replaces the CuPy/cuBLAS/cuSOLVER path of
 * reference src/rcabeam/gpu_clutter.py:112-273) ----------------------------
 * iq is a row-major Casorati matrix (rows, frames) of `storage` samples. */

/* out[0] = bits of max(|Re|, |Im|) over n complex values, out[1] = 1 if any
 * component is non-finite, out[2] = bits of 1 / out[0] (1 for all-zero input).
 * accumulate != 0 folds another array into an existing out (max / or), as
 * the reference's joint scale over an RC/CR tuple (gpu_clutter.py:141). */
rcabeam_status rcabeam_iq_absmax_device(const void* iq, int storage, size_t n, unsigned* out, int accumulate);

/* Workspace (bytes) and real-Gram size (doubles) for rcabeam_temporal_gram_device. */
size_t rcabeam_gram_workspace_bytes(size_t rows, size_t frames);
size_t rcabeam_gram_real_elements(size_t frames);

/* Real Gram C = A^T A of A = (iq / scale) viewed as (rows, 2*frames) reals,
 * upper 32x32 blocks only, float64, row-major with leading dimension
 * 32*ceil(2*frames/32); accumulate != 0 adds to real_gram.  complex64 input
 * is divided by *scale (device float, e.g. out + 0 of
 * rcabeam_iq_absmax_device; 0 means 1) and split into bfloat16 hi + lo, all
 * four products on MFMA; FP16/BF16 input is exact in one product and ignores
 * scale.  frames <= 1024. */
rcabeam_status rcabeam_temporal_gram_device(const void* iq, int storage, size_t rows, size_t frames,
                                            const float* scale, double* real_gram, float* workspace,
                                            size_t workspace_bytes, int accumulate);

/* The same Gram over every row_step-th row of iq: rows 0, row_step, ...,
 * (rows - 1) * row_step (rows = number of rows used; workspace as for
 * rcabeam_gram_workspace_bytes(rows, frames)).  No reference counterpart:
 * it forms the temporal basis of the low-rank channel coding from a row
 * sample of the channel Casorati matrix. */
rcabeam_status rcabeam_temporal_gram_sampled_device(const void* iq, int storage, size_t rows, size_t row_step,
                                                    size_t frames, const float* scale, double* real_gram,
                                                    float* workspace, size_t workspace_bytes, int accumulate);

/* Exact real Gram of complex64 input in float64 on v_mfma_f64_16x16x4_f64
 * (products exact, float64 accumulation): the complex128 Gram of reference
 * gpu_clutter.py:140-161 without prescaling (float64 cannot overflow on
 * complex64 products).  Same real_gram layout as above; frames <= 640. */
size_t rcabeam_gram_f64_workspace_bytes(size_t rows, size_t frames);
rcabeam_status rcabeam_temporal_gram_f64_device(const void* iq, size_t rows, size_t frames, double* real_gram,
                                                double* workspace, size_t workspace_bytes, int accumulate);

/* The same exact Gram without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The float64 partial sums live in a library-owned device array of
 * RCABEAM_GRAM_F64_SCRATCH_BYTES on each GPU (held until the process ends)
 * instead of a caller workspace; the call allocates nothing and frees nothing.
 * Same result as rcabeam_temporal_gram_f64_device, bit for bit, while the
 * default row groups fit the array (rows <= 16384 * floor(scratch / (blocks *
 * 2 KiB)), 1,638,400 rows at frames = 200); larger inputs use larger row
 * groups (float64 summation order differs).  frames > 640 returns
 * RCABEAM_ERROR_ARGUMENT.  Calls that overlap on one GPU share the array and
 * race; serialize them. */
#define RCABEAM_GRAM_F64_SCRATCH_BYTES ((size_t)64 << 20)
rcabeam_status rcabeam_temporal_gram_f64_noalloc_device(const void* iq, size_t rows, size_t frames,
                                                        double* real_gram, int accumulate);

/* Complex Hermitian Gram G = scale2 * X^H X (complex128, column-major
 * frames x frames) from the real Gram above (reference gpu_clutter.py:160-161). */
rcabeam_status rcabeam_gram_hermitian_device(const double* real_gram, size_t frames, double scale2, void* gram);

/* Hermitian eigendecomposition in float64 on the GPU: eigenvalues ascending
 * in values[frames], eigenvectors overwrite gram (columns).  Only the upper
 * triangle and the real part of the diagonal are read.  frames <= 256: a
 * one-workgroup Householder reduction to a real tridiagonal matrix (the
 * reflectors accumulated in place) and rocSOLVER zstedc; larger: rocSOLVER
 * zheevd.  workspace_e: frames doubles; info: one device int (0 = ok).  The
 * solver handle is created once per device on first use (the only
 * allocation in this ABI). */
rcabeam_status rcabeam_eigh_device(void* gram, double* values, double* workspace_e, int* info, size_t frames);

/* Filtered power Doppler per row, float64, original frame count as
 * denominator (reference gpu_clutter.py:244-273).  vectors: complex128
 * column-major (frames, r).  subtract != 0: power of x - V V^H x (V = the
 * discarded vectors); subtract == 0: sum_k |<x, v_k>|^2 / frames (V = the
 * retained vectors).  Arithmetic after the load is float64. */
rcabeam_status rcabeam_filtered_power_device(const void* iq, int storage, size_t rows, size_t frames,
                                             const void* vectors, size_t r, int subtract, double* power);

/* Filtered IQ (complex64) with the same projector convention
 * (reference gpu_clutter.py:212-241). */
rcabeam_status rcabeam_temporal_filter_device(const void* iq, int storage, size_t rows, size_t frames,
                                              const void* vectors, size_t r, int subtract, void* out);

/* ---- direct temporal SVD and complex128 filtering (reference
 * src/rcabeam/gpu_clutter.py:43-109, 276-336; src/rcabeam/clutter.py:18-44,
 * 74-109) -------------------------------------------------------------------
 * Matrices are column-major; complex128 flags select complex128 (1) or
 * complex64 (0) elements.  The rocBLAS / rocSOLVER handle is the one
 * rcabeam_eigh_device creates per device. */

/* Rows [offset, offset + rows) of a column-major matrix (leading dimension
 * ld) from a row-major (rows, frames) Casorati array (complex64 or
 * complex128), converting the element type; stacks RC/CR arrays by offset. */
rcabeam_status rcabeam_casorati_colmajor_device(const void* iq, int complex128_in, size_t rows, size_t frames,
                                                void* matrix, int complex128_out, size_t ld, size_t offset);

/* SVD of the column-major rows x frames matrix (ld = rows) with rocSOLVER
 * gesvd: singular values descending in values (float64 for complex128, float32
 * for complex64; min(rows, frames) of them), V^H in vh (frames x frames,
 * column-major).  want_u != 0 overwrites the matrix with U (its first
 * min(rows, frames) columns), otherwise the matrix is destroyed.  e_work:
 * frames reals; info: one device int (0 = converged).  reference
 * gpu_clutter.py:84-103 (cuSOLVER gesvd, jobu 'O' / 'N', jobvt 'A'). */
rcabeam_status rcabeam_svd_device(void* matrix, int complex128, size_t rows, size_t frames, int want_u, void* values,
                                  void* vh, void* e_work, int* info);

/* U[:, k] *= sigma_k for k in [start, stop), U[:, k] = 0 otherwise (rows x
 * frames, leading dimension ld): the retained singular values of reference
 * gpu_clutter.py:299-301 folded into U. */
rcabeam_status rcabeam_scale_columns_device(void* u, int complex128, size_t rows, size_t ld, size_t frames,
                                            const void* values, size_t start, size_t stop);

/* out (rows x frames, row-major) = B A for column-major B (rows x frames,
 * leading dimension ld) and A (frames x frames), in B's precision (rocBLAS
 * gemm), written as complex128 or complex64 (out_complex128).  When the output
 * type differs from B's, workspace holds chunk_rows x frames elements of B's
 * type; chunk_rows also bounds each gemm.  The U Sigma V^H reconstruction
 * (A = V^H) of reference gpu_clutter.py:302-312 and the retained projector
 * (A = V_r V_r^H) of clutter.py:101. */
rcabeam_status rcabeam_times_square_device(const void* b, int complex128, size_t rows, size_t ld, size_t frames,
                                           const void* a, void* out, int out_complex128, void* workspace,
                                           size_t chunk_rows);

/* Complex128 Gram X^H X (upper triangle, frames x frames column-major) of a
 * column-major complex128 matrix with rocBLAS herk: the Gram of reference
 * clutter.py:29-33 for complex128 input. */
rcabeam_status rcabeam_gram_colmajor_device(const void* matrix, size_t rows, size_t ld, size_t frames, void* gram);

/* P = V_r V_r^H (frames x frames, column-major complex128) for the columns
 * [start, stop) of a column-major complex128 basis. */
rcabeam_status rcabeam_projector_device(const void* basis, size_t frames, size_t start, size_t stop, void* projector);

/* ---- RF -> IQ demodulation (no reference counterpart: reference
 * ROADMAP.md:58 lists RF-to-IQ processing as open work) ---------------------
 * Real RF rf (ns_rf, nc, na, nt) of rf_type 0 (float32) or 1 (int16) at fs_rf
 * Hz, first sample at t_start, to complex64 IQ at fs_rf / decimation:
 *   iq[m] = sum_k 2 taps[k] rf[m decimation + k - (ntaps - 1) / 2]
 *           exp(-2j pi f_demod t_{m decimation + k - (ntaps - 1) / 2}),
 * t_n = t_start + n / fs_rf, RF outside [0, ns_rf) zero, so iq sample m sits
 * at t_start + m decimation / fs_rf and an echo delayed by tau comes out as
 * env(t - tau) exp(-2j pi f_demod tau), the convention of the reference
 * simulator and delay-and-sum (src/rcabeam/sim.py:188, 385).  taps: ntaps
 * (odd, <= 255) real low-pass coefficients (device).  ns_iq <= ceil(ns_rf /
 * decimation).  packed == 0: iq in (ns_iq, nc, na, nt); packed != 0: iq in the
 * frame-inner (nc, na, ns_iq, nt) layout of rcabeam_pack_ensemble_device. */
rcabeam_status rcabeam_rf_to_iq_device(const void* rf, int rf_type, size_t ns_rf, size_t nc, size_t na, size_t nt,
                                       const float* taps, size_t ntaps, size_t decimation, float fs_rf,
                                       float f_demod, float t_start, void* iq, size_t ns_iq, int packed);

/* ---- synthetic acquisitions (GPU counterpart of reference
 * src/rcabeam/sim.py:131-189, scaled to full-size scenes) ------------------
 * Gaussian-envelope echoes exp(-((t - tau)/sigma_t)^2) exp(-2j pi fd tau) of
 * static scatterers (x, z, y; complex amplitude; group index into the per-frame
 * complex factors group_factor[nt][n_groups]) and moving scatterers (position
 * per frame moving_pos[nt][n_moving][3]), plus circular complex Gaussian noise
 * of per-component standard deviation noise_std (counter-based, seeded).
 * out: complex64 (ns, nc, na, nt).  config 0 = RC, 1 = CR.  Dynamic LDS is
 * 8*ns*(n_groups + 16) bytes and must stay <= 160 KiB. */
size_t rcabeam_simulate_lds_bytes(size_t ns, int n_groups);
rcabeam_status rcabeam_simulate_channels_device(const double* static_pos, const double* static_amp,
                                                const int* static_group, size_t n_static, const double* group_factor,
                                                int n_groups, const double* moving_pos, const double* moving_amp,
                                                size_t n_moving, const float* elements, const float* angles,
                                                void* out, size_t ns, size_t nc, size_t na, size_t nt, int config,
                                                float t_start, float c, float fs, float fd, float sigma_t,
                                                float noise_std, unsigned long long seed);

/* Complex mean over frames of the XDoppler signal of per-angle volumes:
 * out[2v] + i out[2v + 1] = mean_t (sum_{rc span} iq) conj(sum_{cr span} iq);
 * arguments as rcabeam_xdoppler_pd_device, whose output is its modulus
 * (reference xdoppler.py:65-117 output="real" / "complex", computed there in
 * NumPy). */
rcabeam_status rcabeam_xdoppler_mean_device(const void* iq, float* out, size_t n_voxels, size_t n_angles,
                                            size_t n_frames, size_t rc_start, size_t rc_count, size_t cr_start,
                                            size_t cr_count);

/* ---- Angle-by-angle combiners on a grid.  RC-FMAS (method 2) and St-SW
 * (method 4) need each angle's channel sum per aperture (reference
 * src/rcabeam/ensemble.cu:236-315 with METHOD 2, 459-590): run
 * rcabeam_opw_grid_device once per angle with method 5 (one angle, channels
 * packed angle-major) and fold each angle's (nv, 2, nt) separate-aperture IQ
 * into an accumulator with rcabeam_angle_accumulate_device (angles in
 * increasing order starting at 0; the first touch of a field stores, so the
 * accumulator needs no clearing), then reduce with
 * rcabeam_angle_combine_finish_device. */

/* float32 accumulator fields per (voxel, frame): 4 for RC-FMAS, 4 + 6 k for
 * St-SW with k subsets (2 <= k <= 4); 0 for unsupported arguments.  The
 * accumulator holds (nv, fields, nt) float32. */
size_t rcabeam_angle_accumulator_fields(int method, int k);

/* Fold angles first_angle .. first_angle + count - 1 (count <= 8; their
 * separate-aperture IQ (nv, 2, nt) complex64, `stride` complex64 apart) into
 * acc in one pass: RC-FMAS sums sgnsqrt(RC), sgnsqrt(CR) (reference
 * fmas.py:31-51); St-SW sums RC, CR and, for subset angle % k, RC, CR,
 * |RC|^2, |CR|^2 (reference ensemble.cu:487-497). */
rcabeam_status rcabeam_angle_accumulate_device(const void* apertures, size_t count, size_t stride, size_t nv,
                                               size_t nt, int method, int k, int first_angle, void* acc);

/* Reduce the accumulator over frames.  RC-FMAS: signal (optional, (nv, nt)
 * complex64) = A B, pd[2v] = mean |A B|^2 (reference ensemble.cu:261-275).
 * St-SW over na angles: pd[2v] = |mean R conj(Q)|, pd[2v + 1] = its imaginary
 * part, weights[v] = the unnormalised St-SW weight (reference
 * ensemble.cu:499-590; the caller divides by the volume maximum as
 * rcabeam_ensemble_device's callers do). */
rcabeam_status rcabeam_angle_combine_finish_device(const void* acc, size_t nv, size_t nt, size_t na, int method,
                                                   int k, float* pd, void* signal, float* weights);

/* DMAS-CCF-ACF on a grid (reference src/rcabeam/dmas.py:32-85, the DMAS kernel
 * of src/rcabeam/ensemble.cu:143-234): the accumulator holds (nv, 9, nt)
 * float32.  Fold in the separate-aperture IQ (nv, 2, nt) of every channel
 * (per_channel = 1: rcabeam_opw_grid_device method 5 on channel i of both
 * apertures, all angles; adds sgnsqrt(RC_i) + sgnsqrt(CR_i) and
 * |RC_i| + |CR_i|) and of every angle (per_channel = 0: one angle, all
 * channels; adds RC_a, |RC_a|^2, CR_a, |CR_a|^2); first = 1 on the first
 * call of each kind stores instead of adding.  Per channel, `count` (<= 8)
 * outputs `stride` complex64 apart are folded in one pass; per angle,
 * count = 1. */
size_t rcabeam_dmas_accumulator_fields(void);
rcabeam_status rcabeam_dmas_accumulate_device(const void* apertures, size_t count, size_t stride, size_t nv,
                                              size_t nt, int per_channel, int first, void* acc);

/* pd[2v] = mean over frames of DMAS x CCF x ACF for the grid (x, z, y) the
 * accumulator covers; the CCF divides by the channels of both apertures
 * inside the dynamic aperture (fn <= 0: all). */
rcabeam_status rcabeam_dmas_finish_device(const void* acc, const float* x, const float* z, const float* y, size_t nx,
                                          size_t nz, size_t ny, const float* xe, const float* ye, size_t nc,
                                          size_t na, size_t nt, float fn, float* pd);

/* ---- Hybrid DMAS-CCF-ACF approximation (reference
 * examples/_support/dmas_delay.cu:21-151, driven by
 * examples/_support/dmas_delay.py:61-154).  Channels are packed complex64
 * (channels, angles, samples, frames); tables are float32 quadruples.
 *
 * Angle planes (dmas_delay.cu:96-114): planes (nv, nk, na, nt) complex64,
 * entry (v, k, a) = sum over channels with prx.w != 0 of the sample at
 * prx.x + ptx[k, a] rotated by (prx.y, prx.z); prx (nv, nc, nk, 4), ptx
 * (nk, na).  Follow with rcabeam_two_stage_spline_device (lines nv, npz nk,
 * chunk = count = nt) for the second derivatives. */
rcabeam_status rcabeam_hybrid_planes_device(const void* packed, const float* prx, const float* ptx, void* planes,
                                            size_t nv, size_t nc, size_t na, size_t ns, size_t nt, size_t nk);

/* Channel tables of depth slab `slab` (dmas_delay.cu:21-38): tables
 * (nu, nc, nq, nt) complex64, entry (u, ch, k) = sum over angles of the sample
 * at anchor.x + k step rotated by (anchor.y, anchor.z), anchor =
 * anchors[slab, u, a] of anchors (n_slabs, nu, na, 4); slab < n_slabs is the
 * caller's to keep. */
rcabeam_status rcabeam_hybrid_tables_device(const void* packed, const float* anchors, void* tables, size_t ns,
                                            size_t nc, size_t na, size_t nt, size_t nu, size_t slab, size_t nq,
                                            float step);

/* Power of depths [first, first + depth) (dmas_delay.cu:116-151): for each
 * voxel and frame, channels with rx.x >= 0 (rx (nz, nv, nc, 4); RC receives
 * along x, CR along y) read their table row (u, ch) at rx.x * inverse_step
 * (cubic != 0: Catmull-Rom inside the table), rotated by (rx.y, rx.z), into
 * the signed-root DMAS and CCF sums; each angle's spline value of the planes
 * at tx.x (tx (nz, nu, na, 4)), rotated by (tx.y, tx.z), into the ACF sums;
 * power[(ix nz + iz) ny + iy] = mean over frames of DMAS x CCF x ACF. */
rcabeam_status rcabeam_hybrid_finish_device(const void* tables_rc, const void* tables_cr, const void* planes_rc,
                                            const void* planes_cr, const void* second_rc, const void* second_cr,
                                            const float* rx_rc, const float* tx_rc, const float* rx_cr,
                                            const float* tx_cr, float* power, size_t nx, size_t nz, size_t ny,
                                            size_t nc, size_t na, size_t nt, size_t first, size_t depth, size_t nq,
                                            size_t nk, float inverse_step, int cubic);

/* ---- Two-stage OPW approximation (reference src/rcabeam/two_stage.py,
 * src/rcabeam/two_stage.cu).  Each aperture is beamformed onto reference
 * planes (receive line x reference depth, one plane per elevation slab), a
 * not-a-knot cubic spline is solved along depth, and every voxel is expanded
 * from it after mapping its elevation offset to a depth shift
 * offset * tan(angle / 2).  Planes are (n_planes * nl, knots, na, chunk)
 * complex64 with the line index plane * nl + line. */

/* Reference planes of one aperture over one depth band (reference kernel
 * `planes`, two_stage.cu:20-84): out[line, k, angle, f] for k < npz, the sum
 * over receive channels with |lateral - element| <= aperture_radius (all
 * channels when aperture_radius < 0) of the linearly interpolated sample at
 * the delay of reference depth knots[knot_first + k], rotated by
 * e^{+2 pi i fd tau} (baseband = 0) or by the residual e^{+2 pi i fd (r - z)/c}
 * of the closest-element carrier (baseband = 1).  Knots that also lie in the
 * previous band [previous_first, previous_first + previous_count) of a line
 * whose same_aperture[band_id * nl + line % nl] is set are copied from
 * `previous` (same layout, previous_count knots).  raw: (ns, nc, na, nt)
 * complex64, frames first .. first + count - 1 of it into chunk slots. */
rcabeam_status rcabeam_two_stage_planes_device(const void* raw, const float* lateral, const float* elements,
                                               const float* angles, const float* starts, const float* anchors,
                                               const float* knots, const unsigned char* same_aperture, void* out,
                                               const void* previous, int band_id, int knot_first, int previous_first,
                                               int previous_count, size_t ns, size_t nc, size_t na, size_t nt,
                                               size_t nl, size_t n_planes, size_t npz, size_t chunk, size_t first,
                                               size_t count, float c, float fs, float fd, float aperture_radius,
                                               int baseband);

/* Not-a-knot second derivatives times the knot step squared along the knot
 * axis of `planes` (reference kernel `spline`, two_stage.cu:86-111; NumPy
 * two_stage.py:119-143).  upper / inverse: Thomas factors u[1] = 0,
 * inv[i] = 1 / (4 - u[i-1]), u[i] = inv[i] for 2 <= i < npz - 2.  npz >= 4. */
rcabeam_status rcabeam_two_stage_spline_device(const void* planes, void* second, const float* upper,
                                               const float* inverse, size_t lines, size_t npz, size_t na,
                                               size_t chunk, size_t count);

/* Voxel expansion of depth rows depth_start .. depth_start + depth_count - 1
 * (reference kernel `expand`, two_stage.cu:155-243).  stsw = 0: out is IQ
 * (nx, nz, ny, nt) complex64, frames first .. first + count - 1 written.
 * stsw = 1: out is (10, nx * nz * ny) float32 slow-time sums (stored when
 * first == 0, added otherwise): XDoppler product, then for the subset pairs
 * (00, 11) and (01, 10) the cross-coherence product and both energies.
 * x_phase (nx, na), z_phase (nz, na), y_phase (ny, na): complex64 carrier
 * e^{i (x|z|y) k_(u|z|u)} restored after the lookup (all three or none). */
rcabeam_status rcabeam_two_stage_expand_device(const void* rc_planes, const void* cr_planes, const void* rc_second,
                                               const void* cr_second, const float* z, const float* slopes, void* out,
                                               int stsw, size_t nx, size_t nz, size_t ny, size_t npz, size_t na,
                                               size_t nt, size_t chunk, size_t first, size_t count,
                                               size_t depth_start, size_t depth_count, float plane_z0,
                                               float plane_dz, const int* x_plane, const float* x_offset,
                                               const int* y_plane, const float* y_offset, size_t n_planes,
                                               const void* x_phase, const void* z_phase, const void* y_phase);

/* St-SW weights and power from the sums of rcabeam_two_stage_expand_device
 * (reference kernel `finish_stsw`, two_stage.cu:138-152, and the
 * normalisation of two_stage.py:384-392): weights = |sum of the two subset-pair
 * correlations| / max, power = |mean XDoppler| * weights.  peak: one device
 * word of workspace. */
rcabeam_status rcabeam_two_stage_stsw_finish_device(const float* stats, float* weights, float* power, size_t nv,
                                                    size_t nt, unsigned* peak);

/* Everything rcabeam_two_stage_reconstruct_device needs besides the data. */
typedef struct rcabeam_two_stage_plan {
    size_t ns, nc, na, nt;        /* channel ensemble (ns, nc, na, nt) */
    size_t nx, nz, ny;            /* grid */
    size_t npz;                   /* reference depths */
    size_t n_planes;              /* elevation slabs per aperture */
    size_t chunk;                 /* frames per pass */
    size_t n_bands;               /* depth bands sharing one receive mask */
    /* host arrays, n_bands entries */
    const int* band_depth_start;
    const int* band_depth_count;
    const int* band_knot_first;
    const int* band_knot_count;
    const float* band_radius;     /* aperture radius at the band's first depth (< 0: full) */
    const float* knots_host;      /* npz reference depths */
    /* device arrays */
    const float* x;               /* nx */
    const float* z;               /* nz */
    const float* y;               /* ny */
    const float* x_el;            /* nc */
    const float* y_el;            /* nc */
    const float* angles;          /* na */
    const float* starts;          /* na */
    const float* slopes;          /* na: tan(angle / 2) */
    const float* knots;           /* npz */
    const float* x_centers;       /* n_planes slabs along x (CR transmit anchors) */
    const int* x_plane;           /* nx: slab of each x */
    const float* x_offset;        /* nx: x minus its slab centre */
    const float* y_centers;       /* n_planes slabs along y (RC transmit anchors) */
    const int* y_plane;           /* ny */
    const float* y_offset;        /* ny */
    const unsigned char* x_same;  /* (n_bands, nx) unchanged RC receive sets */
    const unsigned char* y_same;  /* (n_bands, ny) unchanged CR receive sets */
    const float* spline_upper;    /* >= max band knot count */
    const float* spline_inverse;
    const void* x_phase;          /* baseband carrier tables (see the expansion) */
    const void* z_phase;
    const void* y_phase;
    float c, fs, fd, dz;
    int baseband;                 /* remove the closest-element carrier before the spline */
    int stsw;                     /* 0: IQ out; 1: St-SW sums out */
} rcabeam_two_stage_plan;

/* complex64 elements of each of the four plane buffers the reconstruction
 * uses (RC and CR planes and second derivatives, reused across bands). */
size_t rcabeam_two_stage_buffer_elements(const rcabeam_two_stage_plan* plan);

/* The whole two-stage reconstruction (reference
 * TwoStageOpwWorkspace.reconstruct, two_stage.py:306-396): for each frame
 * chunk and depth band, planes and splines of both apertures, then the voxel
 * expansion.  buffers: 4 * rcabeam_two_stage_buffer_elements complex64;
 * out: IQ (nx, nz, ny, nt) complex64, or the (10, nv) St-SW sums (follow with
 * rcabeam_two_stage_stsw_finish_device). */
rcabeam_status rcabeam_two_stage_reconstruct_device(const rcabeam_two_stage_plan* plan, const void* rc,
                                                    const void* cr, void* buffers, void* out);

#ifdef __cplusplus
}
#endif
