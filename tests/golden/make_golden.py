"""Write the golden fixtures in this directory from the reference's own NumPy code.

Run once, in a container where the reference repository is mounted at
/root/reference (it never travels with this repository):

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py two_stage  # only two_stage.npz

Only functions that no closed form, known-answer test or invariant pins
exactly are evaluated here (the reference's exact complex64 delay-and-sum,
the scatterer simulator, the DMAS/St-SW combiners and the clutter filter).
The fixtures hold inputs and the reference's outputs only, at small sizes.
All geometry values are float32-representable, so the reference's float64
arithmetic and the float32 C ABI see identical numbers.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

REFERENCE_SRC = Path("/root/reference/src")
HERE = Path(__file__).resolve().parent


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def main() -> None:
    sys.path.insert(0, str(REFERENCE_SRC))
    from rcabeam import clutter, dmas, sim, stsw  # noqa: PLC0415

    rng = np.random.default_rng(2024)

    # --- delay-and-sum: both apertures, per-angle starts, with and without f-number.
    ns, nc, na, nt = 64, 5, 3, 4
    iq = (rng.normal(size=(ns, nc, na, nt)) + 1j * rng.normal(size=(ns, nc, na, nt))).astype(np.complex64)
    x_el = f32((np.arange(nc) - (nc - 1) / 2) * 200e-6)
    y_el = f32((np.arange(nc) - (nc - 1) / 2) * 180e-6 + 30e-6)
    angles = f32(np.deg2rad([-7.0, 0.0, 4.0]))
    starts = f32([4.0e-6, 4.1e-6, 4.2e-6])
    grid = (f32(np.linspace(-0.6e-3, 0.6e-3, 3)), f32(np.linspace(3.5e-3, 5.8e-3, 5)), f32(np.linspace(-0.4e-3, 0.5e-3, 4)))
    fixtures = {"iq": iq, "x_el": x_el, "y_el": y_el, "angles": angles, "starts": starts}
    fixtures.update({"grid_x": grid[0], "grid_z": grid[1], "grid_y": grid[2]})
    for fn_name, fnumber in (("f3", 3.0), ("full", None)):
        geom = sim.RCAGeometry(x_el, y_el, c=1540.0, fs=20e6, f_demod=15e6, fnumber=fnumber)
        for config in ("RC", "CR"):
            per_channel, per_angle, counts = sim._delay_numpy(iq, angles, starts, grid, geom, config, channels_out=True)
            key = f"{config}_{fn_name}"
            fixtures[f"{key}_channels"] = per_channel
            fixtures[f"{key}_angles"] = per_angle
            fixtures[f"{key}_counts"] = counts
    np.savez_compressed(HERE / "delay.npz", **fixtures)

    # --- scatterer simulator: Gaussian envelope and band-limited pulse.
    geom = sim.RCAGeometry(x_el, y_el, c=1540.0, fs=20e6, f_demod=15e6)
    point = (0.1e-3, 4.2e-3, -0.2e-3)
    sim_fix = {"x_el": x_el, "y_el": y_el, "angles": angles, "point": np.array(point)}
    for config in ("RC", "CR"):
        sim_fix[f"{config}_gauss"] = sim.simulate_point(geom, angles, point, 96, 4e-6, config)
        sim_fix[f"{config}_band"] = sim.simulate_point(geom, angles, point, 96, 4e-6, config, bandwidth_hz=15e6)
    sim_fix["pulse_delays"] = np.array([5.3e-6, 5.55e-6])
    sim_fix["pulse"] = sim.sample_bandlimited_pulse(sim_fix["pulse_delays"], 128, 4e-6, 20e6, 15e6)
    np.savez_compressed(HERE / "simulate.npz", **sim_fix)

    # --- combiners: DMAS-CCF-ACF core and St-SW weights.
    comb = {}
    s_rc = (rng.standard_normal((3, 4, 6)) + 1j * rng.standard_normal((3, 4, 6))).astype(np.complex64)
    s_cr = (rng.standard_normal((3, 4, 6)) + 1j * rng.standard_normal((3, 4, 6))).astype(np.complex64)
    a_rc = (rng.standard_normal((3, 4, 5)) + 1j * rng.standard_normal((3, 4, 5))).astype(np.complex64)
    a_cr = (rng.standard_normal((3, 4, 5)) + 1j * rng.standard_normal((3, 4, 5))).astype(np.complex64)
    norm = rng.integers(1, 12, size=(3, 4)).astype(np.float32)
    comb.update(s_rc=s_rc, s_cr=s_cr, a_rc=a_rc, a_cr=a_cr, norm=norm)
    y, wc, wa = dmas.dmas_ccf_acf_core(s_rc, s_cr, a_rc, a_cr, ccf_norm=norm)
    comb.update(dmas_y=y, dmas_ccf=wc, dmas_acf=wa)
    iq5 = (rng.standard_normal((3, 4, 2, 8, 5)) + 1j * rng.standard_normal((3, 4, 2, 8, 5))).astype(np.complex64)
    comb["stsw_iq"] = iq5
    for k in (2, 3, 4):
        comb[f"stsw_k{k}"] = stsw.st_sw_weights(iq5, np.arange(4), np.arange(4, 8), k=k, z_chunk=2)
    np.savez_compressed(HERE / "combiners.npz", **comb)

    # --- clutter filter: direct SVD and Gram-EVD on a small random matrix.
    x = rng.normal(size=(40, 9)) + 1j * rng.normal(size=(40, 9))
    x[:, :] += 30 * np.outer(rng.normal(size=40), np.exp(1j * np.linspace(0, 1, 9)))
    clut = {"x": x}
    for method in ("svd", "gram"):
        values, _ = clutter.temporal_basis(x, method=method)
        clut[f"{method}_values"] = values
        clut[f"{method}_filtered"] = clutter.svd_clutter_filter(x, rank_start=1, rank_stop=6, method=method)
    np.savez_compressed(HERE / "clutter.npz", **clut)
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))


def two_stage() -> None:
    """Reference `two_stage_opw_numpy` (src/rcabeam/two_stage.py:146-192) at steered angles: one and two
    reference planes, RF and baseband planes, an f-number of 8 and the full aperture, on a small random
    ensemble (its own seed, so the other fixtures are unaffected).  Written to two_stage.npz."""
    sys.path.insert(0, str(REFERENCE_SRC))
    from rcabeam import sim, two_stage as ts  # noqa: PLC0415

    rng = np.random.default_rng(77)
    ns, nc, na, nt = 80, 4, 3, 3
    rc = (rng.normal(size=(ns, nc, na, nt)) + 1j * rng.normal(size=(ns, nc, na, nt))).astype(np.complex64)
    cr = (rng.normal(size=(ns, nc, na, nt)) + 1j * rng.normal(size=(ns, nc, na, nt))).astype(np.complex64)
    x_el = f32((np.arange(nc) - (nc - 1) / 2) * 200e-6)
    y_el = f32((np.arange(nc) - (nc - 1) / 2) * 180e-6 + 20e-6)
    angles = f32([-0.12, 0.03, 0.1])
    starts = f32([4.9e-6, 5.0e-6, 5.1e-6])
    grid = (f32(np.linspace(-0.3e-3, 0.3e-3, 3)), f32(np.linspace(4.0e-3, 4.4e-3, 6)), f32(np.linspace(-0.25e-3, 0.35e-3, 4)))
    fix = {"rc": rc, "cr": cr, "x_el": x_el, "y_el": y_el, "angles": angles, "starts": starts,
           "grid_x": grid[0], "grid_z": grid[1], "grid_y": grid[2]}
    # F = 8 gives a 0.25 mm receive radius at 4 mm: the aperture bands cut through the element row.
    for fn_name, fnumber in (("f8", 8.0), ("full", None)):
        geom = sim.RCAGeometry(x_el, y_el, c=1540.0, fs=20e6, f_demod=15e6, fnumber=fnumber)
        for n_planes in (1, 2):
            for baseband in (False, True):
                key = f"{fn_name}_p{n_planes}_{'bb' if baseband else 'rf'}"
                fix[key] = ts.two_stage_opw_numpy(rc, cr, angles, starts, grid, geom, samples_per_wavelength=8,
                                                  n_planes=n_planes, baseband=baseband)
    np.savez_compressed(HERE / "two_stage.npz", **fix)
    print("wrote two_stage.npz")


if __name__ == "__main__":
    if sys.argv[1:] == ["two_stage"]:
        two_stage()
    else:
        main()
        two_stage()
