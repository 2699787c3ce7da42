"""Clutter-filtered XDoppler (the recipe of reference README.md:33) against the float64 oracle.

Separately compounded RC and CR IQ, one temporal basis shared by both apertures (the reference's RC/CR
tuple input, gpu_clutter.py:122-127), the singular-vector interval kept in both, |mean_t RC conj(CR)|.
The oracle does the same in float64 NumPy: its per-aperture delay-and-sum, the Gram of the stacked
Casorati rows, the projector onto the kept basis columns.
"""

import numpy as np
import pytest
from conftest import rel_l2

import rcabeam_amd as rb
from oracle import clutter, das, scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rank_start,rank_stop", [(1, None), (1, 12), (0, None)])
def test_filtered_xdoppler_matches_oracle(rank_start, rank_stop):
    print(f"case: filtered XDoppler rank [{rank_start}, {rank_stop})", flush=True)
    acq = scene.point_scene(grid_shape=(16, 24, 12), frames=16)
    geom = rb.RCAGeometry(acq.x_el, acq.y_el, c=acq.c, fs=acq.fs, f_demod=acq.fd, fnumber=acq.fnumber)
    got = rb.filtered_xdoppler_pd(acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, geom, rank_start=rank_start,
                                  rank_stop=rank_stop)
    scan = das.scan_coords(acq.grid)
    (ra, _, _), (qa, _, _) = das.delay_apertures(acq.rc, acq.cr, scan, acq.x_el, acq.y_el, acq.angles, acq.t_start,
                                                 acq.c, acq.fs, acq.fd, acq.fnumber)
    r = ra.sum(axis=1).astype(np.complex128)
    q = qa.sum(axis=1).astype(np.complex128)
    _, basis = clutter.temporal_basis(np.concatenate([r, q]), method="gram")
    start, stop = clutter.rank_interval(r.shape[-1], rank_start, rank_stop)
    kept = basis[:, start:stop]
    fr = (r @ kept) @ kept.conj().T
    fq = (q @ kept) @ kept.conj().T
    expected = np.abs(np.mean(fr * fq.conj(), axis=-1)).reshape(got.shape)
    assert got.shape == tuple(len(a) for a in acq.grid) and got.dtype == np.float32
    err = rel_l2(got, expected)
    print(f"filtered XDoppler rel L2 {err:.2e}", flush=True)
    assert err < 1e-5  # measured 7e-8 to 2e-7
    if rank_start == 0 and rank_stop is None:  # nothing removed: plain XDoppler
        plain = rb.ensemble_pd_from_channels(acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, geom,
                                             method="xdoppler")
        assert rel_l2(got, plain) < 1e-5


@pytest.mark.parametrize("rank_start,rank_stop", [(2, None), (1, 10)])
def test_channel_domain_filter_then_beamform(rank_start, rank_stop):
    """Channel-domain clutter filtering (reference examples/_support/clutter_pipeline.py:87-90): the
    (RC, CR) channel tuple `(samples, channels, angles, frames)` shares one temporal basis over all
    channel rows of both apertures, the interval is kept, and the filtered channels are beamformed.
    Oracle: float64 SVD of the stacked channel rows, then its delay-and-sum and power Doppler.
    Bars: filtered channels rel L2 1e-6 (complex64 output of a complex128 projector), power 1e-5."""
    from oracle import combiners

    print(f"case: channel-domain filter rank [{rank_start}, {rank_stop})", flush=True)
    acq = scene.point_scene(grid_shape=(8, 12, 8), frames=20)
    geom = rb.RCAGeometry(acq.x_el, acq.y_el, c=acq.c, fs=acq.fs, f_demod=acq.fd, fnumber=acq.fnumber)
    rc_f, cr_f = rb.clutter.svd_clutter_filter((acq.rc, acq.cr), rank_start=rank_start, rank_stop=rank_stop)
    assert rc_f.shape == acq.rc.shape and cr_f.dtype == np.complex64
    nt = acq.rc.shape[-1]
    rows = np.concatenate([acq.rc.reshape(-1, nt), acq.cr.reshape(-1, nt)])
    kept = clutter.svd_clutter_filter(rows, rank_start=rank_start, rank_stop=rank_stop)
    exp_rc, exp_cr = kept[: len(rows) // 2].reshape(acq.rc.shape), kept[len(rows) // 2 :].reshape(acq.cr.shape)
    assert rel_l2(rc_f.ravel(), exp_rc.ravel()) < 1e-6
    assert rel_l2(cr_f.ravel(), exp_cr.ravel()) < 1e-6
    pd = rb.ensemble_pd_from_channels(rc_f, cr_f, acq.angles, acq.t_start, acq.grid, geom)
    (ra, _, _), (qa, _, _) = das.delay_apertures(exp_rc, exp_cr, das.scan_coords(acq.grid), acq.x_el, acq.y_el,
                                                 acq.angles, acq.t_start, acq.c, acq.fs, acq.fd, acq.fnumber)
    assert rel_l2(pd.ravel(), combiners.power_doppler(ra.sum(axis=1) + qa.sum(axis=1))) < 1e-5
