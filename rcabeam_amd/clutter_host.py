"""Global temporal clutter filter with the NumPy signature of reference src/rcabeam/clutter.py, computed on MI355X.

The package-level `svd_clutter_filter` and `temporal_basis` are these (as the reference package exports
its NumPy `rcabeam.clutter` functions, src/rcabeam/__init__.py:3): NumPy complex64 / complex128 arrays in,
NumPy arrays out, method "svd" by default, complex128 working precision, the same validation messages
(clutter.py:10-15) and errors.  The arithmetic runs on the GPU:

* "svd": rocSOLVER zgesvd of a column-major complex128 copy of the Casorati matrix (clutter.py:20-26),
  and the filter reconstructs U Sigma V^H of the retained interval with a rocBLAS zgemm (:97-100);
* "gram": complex128 Gram (complex64 input: the exact float64 MFMA Gram, whose products of complex64
  values are exact; complex128 input: rocBLAS zherk), float64 EVD, the reference's negative-eigenvalue
  tolerance 64 eps T max|lambda| (:27-37), and the filter (X V_r) V_r^H as X P with P = V_r V_r^H (:101-103).

The filtered IQ keeps the input shape and dtype (:104).  The GPU-resident path with the reference's CuPy
signatures (`src/rcabeam/gpu_clutter.py`) is `rcabeam_amd.clutter`.
"""

from __future__ import annotations

from typing import Literal

import numpy as np

from rcabeam_amd import _device as dev
from rcabeam_amd import _lib
from rcabeam_amd import clutter as gpu

Solver = Literal["svd", "gram"]


def _matrix(iq) -> np.ndarray:
    """reference clutter.py:10-15."""
    if not isinstance(iq, np.ndarray) or iq.dtype not in (np.complex64, np.complex128):
        raise ValueError("IQ must be a complex64 or complex128 NumPy array")
    if iq.ndim == 0 or iq.size == 0 or not np.isfinite(iq).all():
        raise ValueError("IQ must be nonempty and finite, with frames on the last axis")
    return iq.reshape(-1, iq.shape[-1])


def _device_rows(matrix: np.ndarray):
    t = dev.torch()
    return dev.to_device(matrix, t.complex128 if matrix.dtype == np.complex128 else t.complex64)


def _gram_basis(x):
    """Complex128 Gram + float64 EVD of the device Casorati rows `x` (rows, T): (values desc, basis)."""
    t = dev.torch()
    rows, frames = x.shape
    lib = _lib.load()
    gram = dev.empty((frames, frames), t.complex128)
    if x.dtype == t.complex64:
        real = dev.empty((int(lib.rcabeam_gram_real_elements(frames)),), t.float64)
        nbytes = int(lib.rcabeam_gram_f64_workspace_bytes(rows, frames))
        if nbytes == 0:
            raise ValueError("Frame count out of range for the float64 GPU Gram (1 <= T <= 640)")
        work = dev.empty((nbytes // 8,), t.float64)
        _lib.call("rcabeam_temporal_gram_f64_device", dev.ptr(x), rows, frames, dev.ptr(real), dev.ptr(work), nbytes, 0)
        _lib.call("rcabeam_gram_hermitian_device", dev.ptr(real), frames, 1.0, dev.ptr(gram))
    else:
        matrix, _ = gpu.colmajor_matrix((x,), complex128=True)
        _lib.call("rcabeam_gram_colmajor_device", dev.ptr(matrix), rows, rows, frames, dev.ptr(gram))
    if not bool(t.isfinite(t.view_as_real(gram)).all()):
        raise np.linalg.LinAlgError("Nonfinite Gram matrix; rescale the IQ input")
    values = dev.empty((frames,), t.float64)
    e_work = dev.empty((frames,), t.float64)
    info = dev.empty((1,), t.int32)
    _lib.call("rcabeam_eigh_device", dev.ptr(gram), dev.ptr(values), dev.ptr(e_work), dev.ptr(info), frames)
    if int(info.item()) != 0:
        raise np.linalg.LinAlgError("Eigenvalues did not converge")
    eigenvalues = dev.host(values)
    tolerance = 64 * np.finfo(np.float64).eps * frames * np.max(np.abs(eigenvalues))
    if np.any(eigenvalues < -tolerance):
        raise np.linalg.LinAlgError("Gram matrix has significantly negative eigenvalues")
    # Column-major eigenvectors read row-major: row k is vector k (ascending); the basis has them as columns.
    basis = gram.transpose(0, 1).flip(1)
    return np.sqrt(np.maximum(eigenvalues[::-1], 0)), basis


def _basis(matrix: np.ndarray, method: Solver):
    """(values, device basis) of reference clutter.py:18-44."""
    if method == "svd":
        values, basis, _ = gpu._direct_svd((_device_rows(matrix),), None, precision="complex128")
        singular_values, basis = dev.host(values), basis
    elif method == "gram":
        singular_values, basis = _gram_basis(_device_rows(matrix))
    else:
        raise ValueError("Method must be 'svd' or 'gram'")
    t = dev.torch()
    if not np.isfinite(singular_values).all() or not bool(t.isfinite(t.view_as_real(basis)).all()):
        raise np.linalg.LinAlgError("Nonfinite temporal decomposition")
    return singular_values, basis


def temporal_basis(iq: np.ndarray, *, method: Solver = "svd") -> tuple[np.ndarray, np.ndarray]:
    """Descending singular values `(T,)` and full orthonormal temporal vectors V `(T, T)` as columns
    (reference clutter.py:47-62); zero singular values are padded when there are fewer rows than frames."""
    values, basis = _basis(_matrix(iq), method)
    return values, dev.host(basis)


def _rank_interval(frames: int, rank_start, rank_stop):
    return gpu.rank_interval(frames, rank_start, rank_stop)


def svd_clutter_filter(iq: np.ndarray, *, rank_start: int, rank_stop: int | None = None,
                       method: Solver = "svd") -> np.ndarray:
    """Retain the zero-based temporal singular-vector interval [rank_start, rank_stop) (reference
    clutter.py:74-109); the result has the input's shape and dtype and does not alias it."""
    matrix = _matrix(iq)
    frames = matrix.shape[-1]
    rank_start, stop = _rank_interval(frames, rank_start, rank_stop)
    t = dev.torch()
    if method == "svd":
        filtered = gpu.svd_reference(_device_rows(matrix), rank_start=rank_start, rank_stop=stop)[0]
        if filtered.dtype != (t.complex128 if iq.dtype == np.complex128 else t.complex64):
            raise AssertionError("SVD reconstruction returned the wrong dtype")
    elif method == "gram":
        x = _device_rows(matrix)
        _, basis = _basis(matrix, "gram")
        projector = dev.empty((frames, frames), t.complex128)
        # basis (row-major, vectors as columns) as a column-major matrix is its transpose.
        columns = basis.transpose(0, 1).contiguous()
        _lib.call("rcabeam_projector_device", dev.ptr(columns), frames, rank_start, stop, dev.ptr(projector))
        cm, rows = gpu.colmajor_matrix((x,), complex128=True)
        filtered = dev.empty(tuple(x.shape), x.dtype)
        out_c128 = x.dtype == t.complex128
        chunk = min(1 << 16, rows)
        work = None if out_c128 else dev.empty((chunk, frames), t.complex128)
        _lib.call("rcabeam_times_square_device", dev.ptr(cm), 1, rows, rows, frames, dev.ptr(projector),
                  dev.ptr(filtered), int(out_c128), dev.ptr(work), chunk)
    else:
        raise ValueError("Method must be 'svd' or 'gram'")
    out = dev.host(filtered).reshape(iq.shape)
    if not np.isfinite(out).all():
        raise np.linalg.LinAlgError("Nonfinite filtered IQ; rescale the input")
    return out
