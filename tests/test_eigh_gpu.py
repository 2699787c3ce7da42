"""rcabeam_eigh_device (float64 Hermitian EVD) against NumPy on the matrices the clutter filter meets.

Frames <= 256 take the one-workgroup Householder tridiagonalization with the reflectors accumulated
in place, then rocSOLVER zstedc; larger ones rocSOLVER zheevd.  Eigenvectors are unique only up to a
phase (and within degenerate eigenspaces), so the checks are phase-free: eigenvalues against
numpy.linalg.eigvalsh, the residual ||A V - V diag(w)|| and the orthonormality of V, all relative to
||A||.  Only the upper triangle is referenced (as with zheevd(fill_upper)): the lower one is filled
with garbage.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _eigh_gpu(a):
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib

    n = a.shape[0]
    # Column-major n x n: the row-major view of a.T.
    gram = dev.to_device(np.ascontiguousarray(a.T), torch.complex128)
    values = dev.empty((n,), torch.float64)
    e_work = dev.empty((n,), torch.float64)
    info = dev.empty((1,), torch.int32)
    _lib.call("rcabeam_eigh_device", dev.ptr(gram), dev.ptr(values), dev.ptr(e_work), dev.ptr(info), n)
    assert int(info.item()) == 0
    return values.cpu().numpy(), gram.cpu().numpy().T  # columns = eigenvectors


def _check(a, w, v, tol=1e-12):
    n = a.shape[0]
    scale = max(np.linalg.norm(a, 2), 1e-300)
    assert np.all(np.diff(w) >= -tol * scale)  # ascending
    np.testing.assert_allclose(w, np.linalg.eigvalsh(a), rtol=0, atol=tol * scale * n)
    assert np.linalg.norm(a @ v - v * w[None, :]) <= tol * scale * n
    assert np.abs(v.conj().T @ v - np.eye(n)).max() <= tol * n


def _with_garbage_lower(a, rng):
    g = a.copy()
    low = np.tril_indices(a.shape[0], -1)
    g[low] = rng.normal(size=low[0].size) + 1j * rng.normal(size=low[0].size)
    g[np.diag_indices(a.shape[0])] += 1j * rng.normal(size=a.shape[0])  # imaginary diagonal ignored
    return g


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 200, 256, 300])
def test_random_gram_matches_numpy(n):
    """Gram of a random complex Casorati matrix (the clutter filter's input); 300 takes zheevd."""
    rng = np.random.default_rng(11 + n)
    x = rng.normal(size=(3 * n + 5, n)) + 1j * rng.normal(size=(3 * n + 5, n))
    a = x.conj().T @ x / x.shape[0]
    w, v = _eigh_gpu(_with_garbage_lower(a, rng))
    _check(a, w, v)


def test_clutter_like_spectrum():
    """Three strong modes 60 dB above a flat floor (the full-size scene's shape), T = 200."""
    rng = np.random.default_rng(5)
    n = 200
    q, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    lam = np.concatenate([[1e6, 3e5, 2e5], 1.0 + 0.01 * rng.random(n - 3)])
    a = (q * lam[None, :]) @ q.conj().T
    a = (a + a.conj().T) / 2
    w, v = _eigh_gpu(_with_garbage_lower(a, rng))
    _check(a, w, v)
    # The discarded clutter subspace: projector onto the 3 strongest vectors.
    p = v[:, -3:] @ v[:, -3:].conj().T
    p_ref = q[:, :3] @ q[:, :3].conj().T
    assert np.abs(p - p_ref).max() < 1e-9


def test_diagonal_and_degenerate():
    """Already tridiagonal (every reflector the identity) and a 4-fold degenerate eigenvalue."""
    rng = np.random.default_rng(9)
    d = np.diag(np.arange(1.0, 41.0)).astype(np.complex128)
    w, v = _eigh_gpu(d)
    _check(d, w, v)
    n = 64
    q, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    lam = np.concatenate([np.full(4, 2.0), np.linspace(3.0, 10.0, n - 4)])
    a = (q * lam[None, :]) @ q.conj().T
    a = (a + a.conj().T) / 2
    w, v = _eigh_gpu(_with_garbage_lower(a, rng))
    _check(a, w, v)


def test_zero_matrix():
    a = np.zeros((50, 50), np.complex128)
    w, v = _eigh_gpu(a)
    np.testing.assert_array_equal(w, np.zeros(50))
    assert np.abs(v.conj().T @ v - np.eye(50)).max() <= 1e-12
