"""Low-rank + residual channel coding (rcabeam_amd.lowrank) kernel by kernel, against NumPy float64.

* pack: residual e = c - (c V) V^H rounded to FP16 and the coefficients w = c V as FP16 hi + lo in 2k
  extra frame columns, in the packed (channels, angles, samples, frames + 2k) layout;
* expand: x_t = d_t + sum_j (d_{nt+2j} + d_{nt+2j+1}) conj(V[t, j]) and its mean power;
* round trip: decoding the packed rows themselves (the identity map in place of the delay-and-sum)
  returns the channels to FP16 rounding of the residual only.
"""

import numpy as np
import pytest
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _basis(rng, nt, k):
    q, _ = np.linalg.qr(rng.normal(size=(nt, k)) + 1j * rng.normal(size=(nt, k)))
    return q  # (nt, k), orthonormal columns


def _device_basis(v):
    import torch

    from rcabeam_amd import _device as dev

    return dev.to_device(np.ascontiguousarray(v.T), torch.complex128)  # (k, nt): row j = column j


@pytest.mark.parametrize("nt,k", [(40, 3), (100, 6), (208, 12), (230, 4)])
def test_pack_matches_numpy_coding(nt, k):
    """Every compile-time bound of the pack kernel: k <= 4, 8, 12 and up to 13 or 16 frames per lane."""
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd.lowrank import pack_channels_lowrank

    rng = np.random.default_rng(3 + nt + k)
    ns, nc, na = 9, 5, 3
    v = _basis(rng, nt, k)
    # rows dominated by the basis modes (clutter) plus a small residual
    coef = (rng.normal(size=(ns, nc, na, k)) + 1j * rng.normal(size=(ns, nc, na, k))) * 30
    small = (rng.normal(size=(ns, nc, na, nt)) + 1j * rng.normal(size=(ns, nc, na, nt))) * 0.01
    c = (coef @ v.conj().T + small).astype(np.complex64)
    packed = pack_channels_lowrank(dev.to_device(c, torch.complex64), "float16", _device_basis(v))
    words = packed.cpu().numpy().view(np.float16).reshape(nc, na, ns, nt + 2 * k, 2)
    got = (words[..., 0].astype(np.float64) + 1j * words[..., 1].astype(np.float64)).transpose(2, 0, 1, 3)
    c64 = c.astype(np.complex128)
    w = c64 @ v
    e = c64 - w @ v.conj().T
    e16 = e.real.astype(np.float32).astype(np.float16) + 1j * e.imag.astype(np.float32).astype(np.float16)
    # residual: FP16 rounding of the float64 residual (1 ulp allowed for summation order)
    np.testing.assert_allclose(got[..., :nt], e16, rtol=2 ** -10, atol=2 ** -24)
    whi, wlo = got[..., nt::2], got[..., nt + 1::2]
    assert rel_l2(whi + wlo, w) < 2 ** -20
    assert rel_l2(whi, w) < 2 ** -10


@pytest.mark.parametrize("storage", ["float32", "float16"])
def test_expand_matches_numpy(storage):
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib

    rng = np.random.default_rng(4)
    nv, nt, k = 333, 72, 4
    v = _basis(rng, nt, k)
    coded = (rng.normal(size=(nv, nt + 2 * k)) + 1j * rng.normal(size=(nv, nt + 2 * k))).astype(np.complex64)
    d = coded.astype(np.complex128)
    expected = d[:, :nt] + (d[:, nt::2] + d[:, nt + 1::2]) @ v.conj().T
    code = _lib.STORAGE[storage]
    out = dev.empty((nv, nt), torch.complex64 if code == 0 else torch.int32)
    pd = dev.empty((nv, 2), torch.float32)
    coded_d, basis_d = dev.to_device(coded, torch.complex64), _device_basis(v)  # kept alive across the call
    _lib.call("rcabeam_lowrank_expand_device", dev.ptr(coded_d), dev.ptr(basis_d), k, dev.ptr(out), code,
              dev.ptr(pd), nv, nt)
    x = out if code == 0 else out.view(torch.float16).to(torch.float32).view(torch.complex64)
    x = x.cpu().numpy()
    assert rel_l2(x, expected) < (1e-6 if storage == "float32" else 1e-3)
    assert rel_l2(pd[:, 0].cpu().numpy(), np.mean(np.abs(expected) ** 2, axis=1)) < 1e-6
    np.testing.assert_array_equal(pd[:, 1].cpu().numpy(), 0.0)


def test_round_trip_through_identity_map():
    """Decode the packed rows directly: the identity stands in for the (linear, frame-wise) delay-and-sum."""
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib
    from rcabeam_amd.lowrank import pack_channels_lowrank

    rng = np.random.default_rng(5)
    ns, nc, na, nt, k = 16, 4, 2, 64, 2
    v = _basis(rng, nt, k)
    coef = (rng.normal(size=(ns, nc, na, k)) + 1j * rng.normal(size=(ns, nc, na, k))) * 100
    blood = (rng.normal(size=(ns, nc, na, nt)) + 1j * rng.normal(size=(ns, nc, na, nt))) * 0.01
    c = (coef @ v.conj().T + blood).astype(np.complex64)
    basis = _device_basis(v)
    packed = pack_channels_lowrank(dev.to_device(c, torch.complex64), "float16", basis)
    words = packed.cpu().numpy().view(np.float16).reshape(-1, nt + 2 * k, 2)
    rows = (words[..., 0].astype(np.float32) + 1j * words[..., 1].astype(np.float32)).astype(np.complex64)
    out = dev.empty((rows.shape[0], nt), torch.complex64)
    rows_d = dev.to_device(rows, torch.complex64)
    _lib.call("rcabeam_lowrank_expand_device", dev.ptr(rows_d), dev.ptr(basis), k, dev.ptr(out), 0, 0,
              rows.shape[0], nt)
    back = out.cpu().numpy().reshape(nc, na, ns, nt).transpose(2, 0, 1, 3)
    direct = (c.real.astype(np.float16).astype(np.float32) + 1j * c.imag.astype(np.float16).astype(np.float32))
    weak = np.linalg.norm(blood)
    err_coded = np.linalg.norm(back - c) / weak  # error relative to the weak part
    err_direct = np.linalg.norm(direct - c) / weak
    assert err_coded < 2e-3, err_coded      # FP16 rounding of the residual (~2^-12 of it) and float32 of the rest
    assert err_direct > 50 * err_coded      # rounding the channels directly swamps the weak part


def test_sampled_gram_matches_numpy():
    """rcabeam_temporal_gram_sampled_device on complex64: the real Gram of rows 0, s, 2s, ... (split bfloat16
    products), against float64 NumPy on the same rows; row step 1 equals rcabeam_temporal_gram_device."""
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(11)
    total, nt, step = 5003, 24, 3
    x = (rng.normal(size=(total, nt)) + 1j * rng.normal(size=(total, nt))).astype(np.complex64) * 7.0
    xd = dev.to_device(x, torch.complex64)
    stats = dev.empty((3,), torch.int32)
    _lib.call("rcabeam_iq_absmax_device", dev.ptr(xd), 0, x.size, dev.ptr(stats), 0)
    scale = float(stats.cpu().numpy().view(np.float32)[0])
    ld = int(np.sqrt(int(lib.rcabeam_gram_real_elements(nt))))

    def real_gram(rows, row_step, entry):
        nbytes = int(lib.rcabeam_gram_workspace_bytes(rows, nt))
        work = dev.empty((nbytes // 4,), torch.float32)
        real = dev.empty((ld * ld,), torch.float64)
        if entry == "sampled":
            _lib.call("rcabeam_temporal_gram_sampled_device", dev.ptr(xd), 0, rows, row_step, nt, dev.ptr(stats),
                      dev.ptr(real), dev.ptr(work), nbytes, 0)
        else:
            _lib.call("rcabeam_temporal_gram_device", dev.ptr(xd), 0, rows, nt, dev.ptr(stats), dev.ptr(real),
                      dev.ptr(work), nbytes, 0)
        c = real.cpu().numpy().reshape(ld, ld)[: 2 * nt, : 2 * nt]
        return np.triu(c) + np.triu(c, 1).T  # upper blocks stored

    rows = (total + step - 1) // step
    sample = x[::step].astype(np.complex128) / scale
    a = np.stack([sample.real, sample.imag], axis=-1).reshape(rows, 2 * nt)
    expect = a.T @ a
    got = real_gram(rows, step, "sampled")
    assert rel_l2(got, expect) < 1e-5  # bfloat16 hi + lo keep 16 significant bits (2^-17 = 7.6e-6 per value)
    # a wrong row map (every row, or rows offset by one) is far off
    wrong = np.stack([x[1::step].real, x[1::step].imag], axis=-1).reshape(-1, 2 * nt).astype(np.float64) / scale
    assert rel_l2(got, wrong.T @ wrong) > 1e-2
    np.testing.assert_array_equal(real_gram(total, 1, "sampled"), real_gram(total, 1, "plain"))


def _subspace(g, k, iterations):
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib

    nt = g.shape[0]
    gd = dev.to_device(np.asfortranarray(g).ravel(order="F"), torch.complex128)  # column-major
    out = dev.empty((k, nt), torch.complex128)
    _lib.call("rcabeam_dominant_subspace_device", dev.ptr(gd), nt, k, iterations, dev.ptr(out))
    return out.cpu().numpy()


@pytest.mark.parametrize("nt,k", [(200, 4), (48, 3), (256, 12)])
def test_dominant_subspace_matches_eigendecomposition(nt, k):
    """rcabeam_dominant_subspace_device spans the k leading eigenvectors of a Hermitian matrix with a gap
    (lambda_{k+1} / lambda_k = 1/30), against numpy.linalg.eigh: projectors agree to 1e-10, the basis is
    orthonormal, and its first vector is close to the leading eigenvector (power iteration on column 0, which
    converges at lambda_2 / lambda_1 = 1/2 per round: an angle of order 2^-12 after 12 rounds)."""
    rng = np.random.default_rng(nt + k)
    q = _basis(rng, nt, nt)
    lam = np.concatenate([1000.0 * 0.5 ** np.arange(k), rng.uniform(0.0, 1000.0 * 0.5 ** (k - 1) / 30, nt - k)])
    g = (q * lam) @ q.conj().T
    g = (g + g.conj().T) / 2
    basis = _subspace(g, k, 12)
    assert np.allclose(basis @ basis.conj().T, np.eye(k), atol=1e-12)
    _, vec = np.linalg.eigh(g)
    top = vec[:, nt - k:]
    p_got = basis.T @ basis.conj()  # sum_j b_j b_j^H
    p_ref = top @ top.conj().T
    assert np.abs(p_got - p_ref).max() < 1e-10
    assert abs(abs(np.vdot(basis[0], vec[:, -1])) - 1.0) < 1e-6
    # a transposed (unconjugated) product would span the conjugate subspace, far from the true one
    p_conj = top.conj() @ top.T
    assert np.abs(p_conj - p_ref).max() > 1e-2


def test_dominant_subspace_rank_deficient_and_zero():
    """A rank-2 matrix with k = 4: the basis stays orthonormal and its span holds the range; a zero matrix
    gives a zero basis."""
    rng = np.random.default_rng(5)
    nt, k = 40, 4
    u = _basis(rng, nt, 2)
    g = (u * np.array([5.0, 2.0])) @ u.conj().T
    basis = _subspace(g, k, 12)
    assert np.allclose(basis @ basis.conj().T, np.eye(k), atol=1e-10)
    assert np.all(np.linalg.norm(basis.conj() @ u, axis=0) > 1 - 1e-10)
    assert not np.any(_subspace(np.zeros((nt, nt), np.complex128), k, 3))


def test_channel_basis_spans_planted_clutter():
    """channel_basis recovers a planted 3-mode temporal subspace of RC + CR channels from a row sample."""
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd.lowrank import channel_basis

    rng = np.random.default_rng(13)
    ns, nc, na, nt, k = 64, 8, 3, 48, 3
    v = _basis(rng, nt, k)
    amp = np.array([300.0, 100.0, 30.0])

    def channels():
        coef = (rng.normal(size=(ns, nc, na, k)) + 1j * rng.normal(size=(ns, nc, na, k))) * amp
        noise = (rng.normal(size=(ns, nc, na, nt)) + 1j * rng.normal(size=(ns, nc, na, nt))) * 0.1
        return dev.to_device((coef @ v.conj().T + noise).astype(np.complex64), torch.complex64)

    rc, cr = channels(), channels()
    basis = channel_basis(rc, cr, k).cpu().numpy()  # (k, nt), row j = column j
    assert np.allclose(basis @ basis.conj().T, np.eye(k), atol=1e-10)
    # principal angles: the projection of the planted columns onto the found span keeps their norm
    kept = np.linalg.norm(basis.conj() @ v, axis=0)
    assert np.all(kept > 1 - 1e-5), kept
    # modes in descending strength: the strongest planted mode is (up to phase) the first row
    assert abs(np.vdot(basis[0], v[:, 0])) > 0.999
