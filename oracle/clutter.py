"""Float64 / complex128 restatement of the temporal SVD clutter filter.

TEST INFRASTRUCTURE ONLY (see `oracle/__init__.py`).

Follows:
* `rank_interval`   src/rcabeam/clutter.py:65-71  zero-based, stop-exclusive [start, stop)
* `temporal_basis`  src/rcabeam/clutter.py:18-62  "svd" (LAPACK) or "gram" (chunked Gram +
  Hermitian EVD, descending, negative-eigenvalue tolerance 64 eps T max|lambda|)
* `svd_clutter_filter` src/rcabeam/clutter.py:74-109
* `scaled_gram`     src/rcabeam/gpu_clutter.py:141-161 (componentwise scaling by
  max(|Re|, |Im|), Gram / rows, Hermitian symmetrisation)
* `filtered_power`  src/rcabeam/gpu_clutter.py:244-273 (power of the filtered signal,
  original frame count as denominator, float64 squares, the smaller projector)

Pinned in `tests/test_oracle.py` by: the known-spectrum construction from QR
factors (reference tests/test_clutter.py:9-36), retain-all identity and
empty-interval zeros (:39-53), Gram/SVD projector equality (:56-67), residual
power without cancellation (reference tests/test_gpu_clutter.py:160-166), and
golden vectors written by the reference `svd_clutter_filter`
(`tests/golden/clutter.npz`).
"""

from __future__ import annotations

import numpy as np

_ROW_CHUNK = 16384


def rank_interval(frames: int, rank_start, rank_stop):
    """Validate `[rank_start, rank_stop)` (clutter.py:65-71)."""
    stop = frames if rank_stop is None else rank_stop
    for v in (rank_start, stop):
        if isinstance(v, (bool, np.bool_)) or not isinstance(v, (int, np.integer)):
            raise ValueError("Rank bounds must be integers")
    if not 0 <= rank_start <= stop <= frames:
        raise ValueError("Require 0 <= rank_start <= rank_stop <= number of frames")
    return int(rank_start), int(stop)


def _matrix(iq: np.ndarray) -> np.ndarray:
    arr = np.asarray(iq)
    if arr.ndim == 0 or arr.size == 0 or not np.iscomplexobj(arr) or not np.isfinite(arr).all():
        raise ValueError("IQ must be a nonempty finite complex array with frames on the last axis")
    return arr.reshape(-1, arr.shape[-1])


def scaled_gram(iq: np.ndarray):
    """Return `(gram, scale, rows)`: X^H X / rows of X / scale (gpu_clutter.py:141-161)."""
    matrix = _matrix(iq).astype(np.complex128)
    scale = float(max(np.abs(matrix.real).max(), np.abs(matrix.imag).max()))
    rows, frames = matrix.shape
    gram = np.zeros((frames, frames), np.complex128)
    for start in range(0, rows, _ROW_CHUNK):
        block = matrix[start : start + _ROW_CHUNK] / (scale or 1.0)
        gram += block.conj().T @ block / rows
    return (gram + gram.conj().T) * 0.5, scale, rows


def temporal_basis(iq: np.ndarray, *, method: str = "svd"):
    """Descending singular values `(T,)` and column basis `(T, T)` (clutter.py:18-62)."""
    matrix = _matrix(iq).astype(np.complex128)
    frames = matrix.shape[-1]
    if method == "svd":
        _, values, vh = np.linalg.svd(matrix, full_matrices=matrix.shape[0] < frames)
        values = np.pad(values, (0, frames - len(values)))
        basis = vh.conj().T
    elif method == "gram":
        gram = np.zeros((frames, frames), np.complex128)
        for start in range(0, len(matrix), _ROW_CHUNK):
            block = matrix[start : start + _ROW_CHUNK]
            gram += block.conj().T @ block
        if not np.isfinite(gram).all():
            raise np.linalg.LinAlgError("Nonfinite Gram matrix; rescale the IQ input")
        eigenvalues, basis = np.linalg.eigh(gram)
        tolerance = 64 * np.finfo(np.float64).eps * frames * np.max(np.abs(eigenvalues))
        if np.any(eigenvalues < -tolerance):
            raise np.linalg.LinAlgError("Gram matrix has significantly negative eigenvalues")
        values = np.sqrt(np.maximum(eigenvalues[::-1], 0))
        basis = basis[:, ::-1]
    else:
        raise ValueError("Method must be 'svd' or 'gram'")
    return values, basis


def svd_clutter_filter(iq: np.ndarray, *, rank_start, rank_stop=None, method: str = "svd") -> np.ndarray:
    """Keep singular-vector interval `[rank_start, rank_stop)` (clutter.py:74-109).

    The result is complex128 (the reference casts back to the input dtype).
    """
    matrix = _matrix(iq).astype(np.complex128)
    start, stop = rank_interval(matrix.shape[-1], rank_start, rank_stop)
    if method == "svd":
        u, values, vh = np.linalg.svd(matrix, full_matrices=False)
        values = values.copy()
        values[:start] = 0
        values[stop:] = 0
        out = (u * values) @ vh
    else:
        _, basis = temporal_basis(matrix, method=method)
        kept = basis[:, start:stop]
        out = (matrix @ kept) @ kept.conj().T
    return out.reshape(np.shape(iq))


def filtered_power(iq: np.ndarray, basis: np.ndarray, *, rank_start, rank_stop=None) -> np.ndarray:
    """Filtered power Doppler, original T denominator (gpu_clutter.py:244-273).

    Uses the retained coordinates when at most half the ranks are kept and the
    residual X - (X V_c) V_c^H of the discarded ones otherwise, squaring in
    float64 so neither path subtracts large energies.
    """
    matrix = _matrix(iq).astype(np.complex128)
    frames = matrix.shape[-1]
    start, stop = rank_interval(frames, rank_start, rank_stop)
    basis = np.asarray(basis, dtype=np.complex128)
    if start == stop:
        return np.zeros(np.shape(iq)[:-1])
    if stop - start > frames - (stop - start):
        discarded = np.concatenate((basis[:, :start], basis[:, stop:]), axis=1)
        values = matrix - (matrix @ discarded) @ discarded.conj().T
    else:
        values = matrix @ basis[:, start:stop]
    power = np.sum(values.real**2 + values.imag**2, axis=-1) / frames
    return power.reshape(np.shape(iq)[:-1])
