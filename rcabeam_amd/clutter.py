"""Global temporal SVD clutter filter on MI355X (Gram on MFMA, float64 EVD on the GPU).

Mirrors reference src/rcabeam/gpu_clutter.py: `temporal_basis` (:112-193,
methods "gram" and "svd"), the direct SVD (:43-109), `apply_temporal_filter`
(:219-241), `filtered_power` (:244-273), `svd_reference` (:276-319) and
`svd_clutter_filter` (:322-336), plus the rank-interval rules of
src/rcabeam/clutter.py:65-71.  Inputs are torch GPU tensors (complex64, or
16-bit packed IQ given with `storage=`), or NumPy arrays, which are uploaded
and answered with NumPy.  A tuple of arrays estimates one shared basis from
all rows (RC/CR).

The direct SVD runs rocSOLVER gesvd on a column-major complex128 (or
complex64) copy of the stacked Casorati rows, as the reference runs cuSOLVER
gesvd; its filter reconstructs U Sigma V^H of the retained interval with a
rocBLAS gemm.

Numerics: precision "complex128" forms the Gram exactly on float64 MFMA
(complex64 products are exact in float64), "complex64" uses split bfloat16
(hi + lo) products on MFMA with short float32 runs reduced in float64, and
16-bit packed IQ is exact in one fp16/bf16 MFMA product.  The T x T
eigendecomposition is float64 (rocSOLVER); the projector and power arithmetic
is float64.  Values are sqrt(lambda) * scale * sqrt(rows) in
descending order and the basis columns follow that order.
"""

from __future__ import annotations

from time import perf_counter
from typing import Literal

import numpy as np

from rcabeam_amd import _device as dev
from rcabeam_amd import _lib

Solver = Literal["gram", "svd"]
# Rows per gemm of the SVD reconstruction (bounds its complex128 staging block).
_GEMM_ROWS = 1 << 16


def rank_interval(frames: int, rank_start, rank_stop):
    """Validate the zero-based, stop-exclusive interval (reference clutter.py:65-71)."""
    stop = frames if rank_stop is None else rank_stop
    for v in (rank_start, stop):
        if isinstance(v, (bool, np.bool_)) or not isinstance(v, (int, np.integer)):
            raise ValueError("Rank bounds must be integers")
    if not 0 <= rank_start <= stop <= frames:
        raise ValueError("Require 0 <= rank_start <= rank_stop <= number of frames")
    return int(rank_start), int(stop)


def _storage_code(x, storage: str | None) -> int:
    t = dev.torch()
    if storage in (None, "complex64", "float32"):
        if x.dtype != t.complex64:
            raise ValueError("IQ must contain complex64 arrays (or 16-bit packed IQ with storage=)")
        return 0
    if storage not in ("float16", "bfloat16"):
        raise ValueError("storage must be complex64, float16 or bfloat16")
    if x.dtype not in (t.int32, t.float32) and x.element_size() != 4:
        raise ValueError("16-bit packed IQ must be a 4-byte-per-sample tensor")
    return _lib.STORAGE[storage]


def _inputs(iq, storage=None, allow_complex128=False):
    arrays = iq if isinstance(iq, tuple) else (iq,)
    if not arrays:
        raise ValueError("At least one IQ array is required")
    host = not all(dev.is_tensor(a) for a in arrays)
    t = dev.torch()
    out = []
    for a in arrays:
        if not dev.is_tensor(a):
            arr = np.asarray(a)
            if arr.dtype not in (np.complex64, np.complex128):
                raise ValueError("IQ must contain complex64 or complex128 arrays")
            if arr.dtype == np.complex128 and not allow_complex128:
                raise ValueError("complex128 IQ is not supported on the GPU Gram path; pass complex64")
            a = dev.to_device(arr, t.complex128 if arr.dtype == np.complex128 else t.complex64)
        elif not a.is_cuda:
            raise ValueError("IQ tensors must be on the GPU")
        if a.ndim == 0 or a.numel() == 0:
            raise ValueError("IQ must be nonempty with frames on the last axis")
        if allow_complex128 and a.dtype == t.complex128 and storage is None:
            pass
        else:
            _storage_code(a, storage)
        out.append(a.contiguous())
    if any(a.shape[-1] != out[0].shape[-1] for a in out):
        raise ValueError("All IQ arrays must share the frame count")
    return tuple(out), host


def _scale_and_check(arrays, code):
    """Joint max(|Re|, |Im|) over all arrays and a finiteness check (one D2H of 12 bytes)."""
    t = dev.torch()
    stats = dev.empty((3,), t.int32)
    for i, a in enumerate(arrays):
        _lib.call("rcabeam_iq_absmax_device", dev.ptr(a), code, a.numel(), dev.ptr(stats), int(i > 0))
    host = stats.cpu().numpy().view(np.uint32)
    if host[1]:
        raise ValueError("IQ must be finite")
    return float(host[:1].view(np.float32)[0]), stats


def _check_finite(arrays):
    """IQ must be finite (reference gpu_clutter.py:38-39): one |max| pass per array, one 12-byte read."""
    t = dev.torch()
    stats = dev.empty((3,), t.int32)
    for i, a in enumerate(arrays):
        if a.dtype == t.complex128:
            if not bool(t.isfinite(t.view_as_real(a)).all()):
                raise ValueError("IQ must be finite")
            continue
        _lib.call("rcabeam_iq_absmax_device", dev.ptr(a), 0, a.numel(), dev.ptr(stats), int(i > 0))
        if stats.cpu().numpy().view(np.uint32)[1]:
            raise ValueError("IQ must be finite")


def colmajor_matrix(arrays, complex128: bool):
    """Column-major copy of the stacked Casorati rows of `arrays` (rows, frames): a `(frames, rows)` tensor
    (gpu_clutter.py:57-65: one Fortran copy also joins RC/CR)."""
    t = dev.torch()
    frames = arrays[0].shape[-1]
    rows = sum(a.numel() // frames for a in arrays)
    matrix = dev.empty((frames, rows), t.complex128 if complex128 else t.complex64)
    offset = 0
    for a in arrays:
        r = a.numel() // frames
        _lib.call("rcabeam_casorati_colmajor_device", dev.ptr(a), int(a.dtype == t.complex128), r, frames,
                  dev.ptr(matrix), int(complex128), rows, offset)
        offset += r
    return matrix, rows


def _direct_svd(arrays, timings, *, spatial: bool = False, precision: str = "complex128"):
    """Singular values `(T,)` (padded with zeros), basis V `(T, T)` and, with `spatial`, the column-major U
    (gpu_clutter.py:43-109)."""
    if precision not in ("complex64", "complex128"):
        raise ValueError("Precision must be complex64 or complex128")
    t = dev.torch()
    single = precision == "complex64"
    frames = arrays[0].shape[-1]
    begin = perf_counter()
    if single and any(a.dtype == t.complex128 for a in arrays):
        for a in arrays:
            if a.dtype == t.complex128 and not bool(t.isfinite(t.view_as_real(a.to(t.complex64))).all()):
                raise np.linalg.LinAlgError("SVD input cannot be represented as finite complex64; rescale the input")
    matrix, rows = colmajor_matrix(arrays, complex128=not single)
    if timings is not None:
        dev.synchronize()
        timings["matrix_seconds"] = perf_counter() - begin
    begin = perf_counter()
    real = t.float32 if single else t.float64
    values = t.zeros((frames,), dtype=real, device=matrix.device)
    vh = dev.empty((frames, frames), t.complex64 if single else t.complex128)  # column-major V^H
    e_work = dev.empty((frames,), real)
    info = dev.empty((1,), t.int32)
    _lib.call("rcabeam_svd_device", dev.ptr(matrix), int(not single), rows, frames, int(spatial), dev.ptr(values),
              dev.ptr(vh), dev.ptr(e_work), dev.ptr(info))
    status = int(info.item())
    if status != 0:
        raise np.linalg.LinAlgError(f"rocSOLVER SVD failed (info={status})")
    if timings is not None:
        timings["decomposition_seconds"] = perf_counter() - begin
    # The column-major V^H read row-major is V^T, so V = conj of that view.
    basis = vh.conj().resolve_conj()
    return values, basis, (matrix, rows) if spatial else None


def temporal_basis(iq, *, method: Solver = "gram", precision: str = "complex128", timings=None, diagnostics=None,
                   storage: str | None = None):
    """Descending singular values `(T,)` and column basis `(T, T)` (gpu_clutter.py:112-193).

    method "gram": Gram on MFMA + float64 EVD (values float64); "svd": direct rocSOLVER gesvd in `precision`
    (complex128 or complex64 input; values float64 or float32)."""
    if precision not in ("complex64", "complex128"):
        raise ValueError("Precision must be complex64 or complex128")
    if method == "svd":
        if storage not in (None, "complex64", "float32"):
            raise ValueError("The direct SVD takes complex64 or complex128 IQ")
        arrays, host = _inputs(iq, None, allow_complex128=True)
        dev.require_default_stream("Clutter filtering requires the default stream")
        _check_finite(arrays)
        values, basis, _ = _direct_svd(arrays, timings, precision=precision)
        t = dev.torch()
        if not bool(t.isfinite(values).all()) or not bool(t.isfinite(t.view_as_real(basis)).all()):
            raise np.linalg.LinAlgError("Nonfinite temporal decomposition")
        if host:
            return dev.host(values), dev.host(basis)
        return values, basis
    if method != "gram":
        raise ValueError("Method must be 'svd' or 'gram'")
    arrays, host = _inputs(iq, storage)
    code = _storage_code(arrays[0], storage)
    t = dev.torch()
    dev.require_default_stream("Clutter filtering requires the default stream")
    frames = arrays[0].shape[-1]
    begin = perf_counter()
    scale, stats = _scale_and_check(arrays, code)
    if code != 0:
        scale = 1.0  # 16-bit products are exact; no prescaling
    rows = sum(a.numel() // frames for a in arrays)
    lib = _lib.load()
    real = dev.empty((int(lib.rcabeam_gram_real_elements(frames)),), t.float64)
    exact = code == 0 and precision == "complex128"
    for i, a in enumerate(arrays):
        r = a.numel() // frames
        if exact:
            nbytes = int(lib.rcabeam_gram_f64_workspace_bytes(r, frames))
            if nbytes == 0:
                raise ValueError("Frame count out of range for the float64 GPU Gram (1 <= T <= 640)")
            work = dev.empty((nbytes // 8,), t.float64)
            _lib.call("rcabeam_temporal_gram_f64_device", dev.ptr(a), r, frames, dev.ptr(real), dev.ptr(work),
                      nbytes, int(i > 0))
        else:
            nbytes = int(lib.rcabeam_gram_workspace_bytes(r, frames))
            if nbytes == 0:
                raise ValueError("Frame count out of range for the GPU Gram (1 <= T <= 1024)")
            work = dev.empty((nbytes // 4,), t.float32)
            _lib.call("rcabeam_temporal_gram_device", dev.ptr(a), code, r, frames, dev.ptr(stats), dev.ptr(real),
                      dev.ptr(work), nbytes, int(i > 0))
    # The float64 path forms the unscaled Gram; fold the 1/scale^2 of the reference's scaled Gram in here.
    scale2 = (1.0 / (scale * scale) if (exact and scale > 0) else 1.0) / rows
    gram = dev.empty((frames, frames), t.complex128)
    _lib.call("rcabeam_gram_hermitian_device", dev.ptr(real), frames, scale2, dev.ptr(gram))
    if timings is not None:
        dev.synchronize()
        timings["matrix_seconds"] = perf_counter() - begin
    begin = perf_counter()
    values = dev.empty((frames,), t.float64)
    e_work = dev.empty((frames,), t.float64)
    info = dev.empty((1,), t.int32)
    _lib.call("rcabeam_eigh_device", dev.ptr(gram), dev.ptr(values), dev.ptr(e_work), dev.ptr(info), frames)
    status = int(info.item())
    if timings is not None:
        timings["decomposition_seconds"] = perf_counter() - begin
    if status != 0:
        raise np.linalg.LinAlgError(f"GPU Hermitian eigensolver failed (info={status})")
    ascending = values.cpu().numpy()
    if not np.isfinite(ascending).all():
        raise np.linalg.LinAlgError("Nonfinite Gram matrix; rescale the IQ input")
    # Negative-eigenvalue tolerance of the reference (gpu_clutter.py:171) at the working precision of the
    # Gram: float64 for the exact Gram, float32 where products or runs are accumulated in float32.
    eps = np.finfo(np.float64 if exact else np.float32).eps
    tolerance = float(64 * eps * frames * np.max(np.abs(ascending)))
    minimum = float(ascending[0])
    if diagnostics is not None:
        diagnostics.update(
            precision=precision,
            input_scale=scale,
            gram_rows=rows,
            minimum_eigenvalue=minimum,
            negative_eigenvalues=int(np.count_nonzero(ascending < 0)),
            negative_tolerance=tolerance,
        )
    if minimum < -tolerance:
        raise np.linalg.LinAlgError(
            f"Gram matrix has significantly negative eigenvalues (minimum={minimum:.6g}, tolerance={tolerance:.6g})"
        )
    out_values = t.sqrt(t.clamp(values.flip(0), min=0)) * (scale * np.sqrt(rows))
    # gram now holds eigenvectors in column-major order: row k of the row-major view is vector k.
    basis = gram.flip(0).transpose(0, 1).contiguous()
    if precision == "complex64":
        basis = basis.to(t.complex64)
    if host:
        return dev.host(out_values), dev.host(basis)
    return out_values, basis


def _vectors(basis, frames: int, start: int, stop: int):
    """The smaller projector's vectors (reference gpu_clutter.py:196-209), complex128 column-major."""
    t = dev.torch()
    b = dev.to_device(basis, t.complex128) if not dev.is_tensor(basis) else basis
    if not b.is_cuda or tuple(b.shape) != (frames, frames) or b.dtype not in (t.complex64, t.complex128):
        raise ValueError("Basis must be a complex (frames, frames) GPU array")
    b = b.to(t.complex128)
    tol = max(1e-12, 32 * float(np.finfo(np.float32 if basis.dtype in (np.complex64, t.complex64) else np.float64).eps))
    if not bool(t.isfinite(t.view_as_real(b)).all()):
        raise ValueError("Basis must be finite and orthonormal")
    gram = b.conj().transpose(0, 1) @ b
    if not bool(t.allclose(gram, t.eye(frames, dtype=t.complex128, device=b.device), rtol=tol, atol=tol)):
        raise ValueError("Basis must be finite and orthonormal")
    subtract = stop - start > frames - (stop - start)
    if subtract:
        cols = list(range(start)) + list(range(stop, frames))
    else:
        cols = list(range(start, stop))
    index = t.tensor(cols, dtype=t.long, device=b.device)
    vectors = b.index_select(1, index).transpose(0, 1).contiguous() if cols else None
    return vectors, len(cols), subtract


def filtered_power(iq, basis, *, rank_start, rank_stop=None, storage: str | None = None):
    """Filtered power Doppler, float64, shape `iq.shape[:-1]` (gpu_clutter.py:244-273)."""
    if isinstance(iq, tuple):
        raise ValueError("Power requires one coherent beamformed IQ array")
    (x,), host = _inputs(iq, storage)
    code = _storage_code(x, storage)
    t = dev.torch()
    frames = x.shape[-1]
    start, stop = rank_interval(frames, rank_start, rank_stop)
    vectors, r, subtract = _vectors(basis, frames, start, stop)
    rows = x.numel() // frames
    power = dev.empty((rows,), t.float64)
    if start == stop:
        power.zero_()
    else:
        _lib.call("rcabeam_filtered_power_device", dev.ptr(x), code, rows, frames, dev.ptr(vectors), r,
                  int(subtract), dev.ptr(power))
    power = power.reshape(tuple(x.shape[:-1]))
    return dev.host(power) if host else power


def apply_temporal_filter(iq, basis, *, rank_start, rank_stop=None):
    """Filtered complex64 IQ with the smaller projector (gpu_clutter.py:219-241)."""
    arrays, host = _inputs(iq)
    t = dev.torch()
    frames = arrays[0].shape[-1]
    start, stop = rank_interval(frames, rank_start, rank_stop)
    vectors, r, subtract = _vectors(basis, frames, start, stop)
    result = []
    for x in arrays:
        rows = x.numel() // frames
        out = dev.empty(tuple(x.shape), t.complex64)
        if not subtract and r == 0:
            out.zero_()
        else:
            _lib.call("rcabeam_temporal_filter_device", dev.ptr(x), 0, rows, frames, dev.ptr(vectors), r,
                      int(subtract), dev.ptr(out))
        result.append(dev.host(out) if host else out)
    return tuple(result) if isinstance(iq, tuple) else result[0]


def svd_reference(iq, *, rank_start, rank_stop=None, timings=None, precision: str = "complex128"):
    """Full direct SVD, discarded singular values zeroed, U Sigma V^H reconstructed (gpu_clutter.py:276-319).

    Returns `(filtered IQ, singular values, V)`; the filtered arrays keep their input dtype.  For a
    tall matrix U overwrites the solver's column-major copy, as in the reference."""
    arrays, host = _inputs(iq, None, allow_complex128=True)
    dev.require_default_stream("Clutter filtering requires the default stream")
    t = dev.torch()
    frames = arrays[0].shape[-1]
    start, stop = rank_interval(frames, rank_start, rank_stop)
    _check_finite(arrays)
    values, basis, (matrix, rows) = _direct_svd(arrays, timings, spatial=True, precision=precision)
    if not bool(t.isfinite(values).all()) or not bool(t.isfinite(t.view_as_real(basis)).all()):
        raise np.linalg.LinAlgError("Nonfinite SVD decomposition")
    begin = perf_counter()
    single = precision == "complex64"
    # Retained singular values folded into U (columns past min(rows, frames) are zeroed), then U' V^H.
    _lib.call("rcabeam_scale_columns_device", dev.ptr(matrix), int(not single), rows, rows, frames, dev.ptr(values),
              start, stop)
    vh = basis.conj().resolve_conj()  # column-major V^H
    result = []
    offset = 0
    for a in arrays:
        r = a.numel() // frames
        out = dev.empty(tuple(a.shape), a.dtype)
        out_c128 = a.dtype == t.complex128
        chunk = min(_GEMM_ROWS, r)
        work = None if out_c128 == (not single) else dev.empty((chunk, frames), t.complex64 if single else t.complex128)
        u_rows = matrix.reshape(-1)[offset:]  # column-major U starting at row `offset`, leading dimension rows
        _lib.call("rcabeam_times_square_device", dev.ptr(u_rows), int(not single), r, rows, frames, dev.ptr(vh),
                  dev.ptr(out), int(out_c128), dev.ptr(work), chunk)
        if not bool(t.isfinite(t.view_as_real(out)).all()):
            raise np.linalg.LinAlgError("Nonfinite SVD reconstruction")
        result.append(out)
        offset += r
    if timings is not None:
        dev.synchronize()
        timings["application_seconds"] = perf_counter() - begin
    filtered = tuple(result) if isinstance(iq, tuple) else result[0]
    if host:
        filtered = tuple(dev.host(o) for o in filtered) if isinstance(filtered, tuple) else dev.host(filtered)
        return filtered, dev.host(values), dev.host(basis)
    return filtered, values, basis


def svd_clutter_filter(iq, *, rank_start, rank_stop=None, method: Solver = "gram", precision: str = "complex128"):
    """Retain the temporal singular-vector interval [rank_start, rank_stop) (gpu_clutter.py:322-336): the direct
    SVD reference for method "svd", the Gram basis and the smaller projector otherwise."""
    if method == "svd":
        return svd_reference(iq, rank_start=rank_start, rank_stop=rank_stop, precision=precision)[0]
    if method != "gram":
        raise ValueError("Method must be 'svd' or 'gram'")
    arrays, _ = _inputs(iq)
    rank_interval(arrays[0].shape[-1], rank_start, rank_stop)
    _, basis = temporal_basis(arrays if isinstance(iq, tuple) else arrays[0], method=method, precision=precision)
    out = apply_temporal_filter(arrays if isinstance(iq, tuple) else arrays[0], basis, rank_start=rank_start,
                                rank_stop=rank_stop)
    host = not all(dev.is_tensor(a) for a in (iq if isinstance(iq, tuple) else (iq,)))
    if host:
        return tuple(dev.host(o) for o in out) if isinstance(out, tuple) else dev.host(out)
    return out
