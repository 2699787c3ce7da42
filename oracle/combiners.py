"""Float64 restatements of the reference combiners (power Doppler and nonlinear beamformers).

TEST INFRASTRUCTURE ONLY (see `oracle/__init__.py`).

Follows:
* `power_doppler`   src/rcabeam/power.py:8-21   mean |x|^2 over the last axis
* `opw_numpy`       src/rcabeam/opw.py:8-21     sum over the angle axis
* `xdoppler_signal` src/rcabeam/xdoppler.py:10-51   (sum RC) * conj(sum CR)
* `signed_sqrt`     src/rcabeam/fmas.py:11-28   x / sqrt(|x|), 0 -> 0
* `rc_fmas`         src/rcabeam/fmas.py:31-51   (sum sgnsqrt RC) * (sum sgnsqrt CR), no conjugate
* `dmas_bruteforce` src/rcabeam/dmas.py:11-29   literal pairwise DMAS (Eq. 6)
* `dmas_ccf_acf_core` src/rcabeam/dmas.py:32-85 factorized DMAS (Eq. 7), CCF and ACF
* `st_sw_weights`   src/rcabeam/stsw.py:10-82   subset-coherence weights
* `ensemble_power`  the per-method outputs of the frame-batched reference binding,
  src/rcabeam/ensemble.py:52-130 with the epilogues of src/rcabeam/ensemble.cu:143-627.

Pinned in `tests/test_oracle.py` by: XDoppler known answer 14 (reference
tests/test_xdoppler.py:35-44), RC-FMAS = literal double sum (reference
tests/test_fmas.py:8-31), DMAS Eq. 6 = Eq. 7 (reference tests/test_dmas.py:17-23),
DMAS linear-scaling known answer DMAS = 2s, CCF = 1/2, ACF = 1 (reference
tests/test_ensemble_targets.py:123-130), St-SW range/normalisation properties
(reference tests/test_stsw.py), and golden vectors written by the reference
`dmas_ccf_acf_core` and `st_sw_weights` (`tests/golden/combiners.npz`).
"""

from __future__ import annotations

import numpy as np


def power_doppler(signal: np.ndarray) -> np.ndarray:
    """Mean squared magnitude over slow time (power.py:8-21)."""
    return np.mean(np.abs(signal) ** 2, axis=-1)


def opw_numpy(iq: np.ndarray) -> np.ndarray:
    """Coherent angle sum over axis -2 (opw.py:8-21)."""
    return np.asarray(iq).sum(axis=-2)


def xdoppler_signal(iq: np.ndarray, rc_idx, cr_idx) -> np.ndarray:
    """s_rc * conj(s_cr) with per-aperture angle sums (xdoppler.py:10-51)."""
    arr = np.asarray(iq)
    return arr[..., rc_idx, :].sum(axis=-2) * np.conj(arr[..., cr_idx, :].sum(axis=-2))


def signed_sqrt(x: np.ndarray) -> np.ndarray:
    """Complex signed square root x / sqrt(|x|), zero at zero (fmas.py:11-28)."""
    arr = np.asarray(x)
    mag = np.abs(arr)
    out = np.zeros_like(arr)
    np.divide(arr, np.sqrt(mag), out=out, where=mag > 0, casting="unsafe")
    return out


def rc_fmas(iq: np.ndarray, rc_idx, cr_idx) -> np.ndarray:
    """RC-FMAS compounded signal (fmas.py:31-51)."""
    arr = np.asarray(iq)
    return signed_sqrt(arr[..., rc_idx, :]).sum(axis=-2) * signed_sqrt(arr[..., cr_idx, :]).sum(axis=-2)


def dmas_bruteforce(s: np.ndarray) -> np.ndarray:
    """Literal pairwise DMAS Eq. 6 over the last axis (dmas.py:11-29)."""
    st = signed_sqrt(s)
    out = np.zeros(s.shape[:-1], dtype=st.dtype)
    for i in range(s.shape[-1] - 1):
        for j in range(i + 1, s.shape[-1]):
            out = out + st[..., i] * np.conj(st[..., j]) + st[..., j] * np.conj(st[..., i])
    return out.real


def dmas_ccf_acf_core(s_rc, s_cr, a_rc, a_cr, *, ccf_norm=None):
    """Factorized DMAS, CCF and ACF terms (dmas.py:32-85)."""
    s = np.concatenate([s_rc, s_cr], axis=-1)
    st = signed_sqrt(s)
    y_dmas = np.abs(st.sum(axis=-1)) ** 2 - np.sum(np.abs(st) ** 2, axis=-1)
    norm = s.shape[-1] if ccf_norm is None else ccf_norm
    with np.errstate(divide="ignore", invalid="ignore"):
        w_ccf = np.maximum(y_dmas, 0) / norm / np.sum(np.abs(s), axis=-1)
    mu_rc = a_rc.mean(axis=-1)
    mu_cr = a_cr.mean(axis=-1)
    var_rc = np.mean(np.abs(a_rc - mu_rc[..., None]) ** 2, axis=-1)
    var_cr = np.mean(np.abs(a_cr - mu_cr[..., None]) ** 2, axis=-1)
    num = np.abs(mu_rc * np.conj(mu_cr))
    den = np.abs(mu_rc) ** 2 + np.abs(mu_cr) ** 2 + (var_rc + var_cr) / 2 - num
    with np.errstate(divide="ignore", invalid="ignore"):
        w_acf = num / den
    return tuple(np.where(np.isfinite(v), v, 0) for v in (y_dmas, w_ccf, w_acf))


def st_sw_weights(iq: np.ndarray, rc_idx, cr_idx, *, k: int = 4, normalize: bool = True) -> np.ndarray:
    """Spatio-temporal similarity weights over `(..., angles, frames)` (stsw.py:10-82).

    With `normalize=False` returns the raw |sum of correlations| before the
    division by the volume maximum (stsw.py:79-82).
    """
    arr = np.asarray(iq).astype(np.complex128)
    rc = arr[..., rc_idx, :]
    cr = arr[..., cr_idx, :]
    m_rc, v_rc, m_cr, v_cr = [], [], [], []
    for subset in range(k):
        for src, means, variances in ((rc, m_rc, v_rc), (cr, m_cr, v_cr)):
            group = src[..., subset::k, :]
            mu = group.mean(axis=-2)
            means.append(mu)
            variances.append(np.mean(np.abs(group - mu[..., None, :]) ** 2, axis=-2))
    ccf, energy = {}, {}
    for ir in range(k):
        for ic in range(k):
            num = m_rc[ir] * np.conj(m_cr[ic])
            den = np.abs(m_rc[ir]) ** 2 + np.abs(m_cr[ic]) ** 2 + (v_rc[ir] + v_cr[ic]) / 2 - np.abs(num)
            with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
                value = num / den
            value[~np.isfinite(value)] = 0
            ccf[ir, ic] = value
            energy[ir, ic] = np.sum(np.abs(value) ** 2, axis=-1)
    acc = np.zeros(arr.shape[:-2], dtype=np.complex128)
    for ir1 in range(k):
        for ic1 in range(k):
            for ir2 in range(ir1 + 1, k):
                for ic2 in range(k):
                    if ic2 == ic1:
                        continue
                    den = np.sqrt(energy[ir1, ic1] * energy[ir2, ic2])
                    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
                        corr = np.sum(ccf[ir1, ic1] * np.conj(ccf[ir2, ic2]), axis=-1) / den
                    corr = np.where(np.isfinite(corr), corr, 0)
                    acc += corr
    weights = np.abs(acc)
    if not normalize:
        return weights
    peak = weights.max()
    return weights / peak if peak > 0 else weights


def ensemble_power(method: str, rc_result, cr_result, *, k: int = 2):
    """Expected `(power, signal)` of the frame-batched ensemble for one method.

    `rc_result` and `cr_result` are `delay_volumes(..., channels_out=True)`
    tuples of one ensemble.  Methods follow src/rcabeam/ensemble.py:52-130 and
    the kernel epilogues of src/rcabeam/ensemble.cu: "opw", "xdoppler",
    "rc_fmas", "dmas", "st_sw" (power only) and "apertures" (signal only).
    Power has shape `(nv,)`; signal `(nv, nt)` (`(nv, 2, nt)` for apertures).
    """
    ra, rch, rn = rc_result
    qa, qch, qn = cr_result
    r, q = ra.sum(axis=1), qa.sum(axis=1)
    if method == "opw":
        signal = r + q
        return power_doppler(signal), signal
    if method == "apertures":
        return np.zeros(len(r)), np.stack([r, q], axis=1)
    if method == "xdoppler":
        signal = r * np.conj(q)
        return np.abs(signal.mean(axis=-1)), signal
    if method == "rc_fmas":
        signal = signed_sqrt(ra).sum(axis=1) * signed_sqrt(qa).sum(axis=1)
        return power_doppler(signal), signal
    if method == "dmas":
        if rch is None or qch is None:
            raise ValueError("DMAS needs per-channel sums (channels_out=True)")
        nt = r.shape[-1]
        frames = []
        for t in range(nt):
            y, wc, wa = dmas_ccf_acf_core(rch[..., t], qch[..., t], ra[..., t], qa[..., t], ccf_norm=rn + qn)
            frames.append(y * wc * wa)
        signal = np.stack(frames, axis=-1)
        return signal.mean(axis=-1), signal
    if method == "st_sw":
        na = ra.shape[1]
        iq = np.concatenate([ra, qa], axis=1)
        weights = st_sw_weights(iq, np.arange(na), np.arange(na, 2 * na), k=k)
        return weights * np.abs((r * np.conj(q)).mean(axis=-1)), None
    raise ValueError(f"Unknown method: {method}")
