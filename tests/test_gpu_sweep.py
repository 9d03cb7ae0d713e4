"""The sweep driver and Green-spectrum reuse on the GPU.

* ``bdda_green_share``: plans of one cuboid share a spectrum; a later
  ``green_init`` on one of them leaves the other's untouched; solves with a
  shared spectrum equal solves with their own bit for bit.
* The sweep driver (``bdda_amd.sweep``, the drop-in of ``run_dda.py``) on the
  reference's stored sweep (``dda_results/pcas_ocbs_simulated_data.hdf5``:
  16 conditions, 100 orientations, block-BiCGSTAB at the reference's
  defaults tol 1e-2 / itermax 25): orientations and r_ve equal the stored
  ones, the Mie values equal the stored ones, every condition matches the
  float64 oracle's solve of it (``sweep_oracle.npz``: S and C within 1e-6 of
  their largest magnitude, or within 10x of the oracle's own rounding
  sensitivity where 100 orientations make the first L x L systems
  ill-conditioned), the stored values at their 2e-2 class (the file
  was written by an earlier code state), one Green spectrum is built per
  (shape, wavelength, medium), and a second run skips every condition.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _npz(name):
    with np.load(os.path.join(GOLDEN, name)) as d:
        return {k: d[k] for k in d.files}


def test_green_share_and_reinit():
    from bdda_amd.plan import Plan
    say("green share")
    a, b = Plan((9, 7, 8)), Plan((9, 7, 8))
    a.green_init(0.05, 9.0)
    ga = a.green_spectrum()
    b.share_green(a)
    assert np.array_equal(b.green_spectrum(), ga)
    b.green_init(0.05, 11.0)  # b gets its own spectrum; a keeps its own
    assert np.array_equal(a.green_spectrum(), ga)
    assert not np.array_equal(b.green_spectrum(), ga)
    c = Plan((9, 7, 9))
    with pytest.raises(ValueError):
        c.share_green(a)


def test_shared_spectrum_solves_equal_fresh_ones():
    from bl_dda.scatterer import DiscreteDipoles, IncidentField, Target
    from mvp_fft.mvp_fft import GreenTensor
    from oracle import gre, orient
    say("shared spectrum across m_p")
    name, n, lf, grid, is_in, _, _ = gre.build_gre_geometry(0.1, 1.2, 1.1, 0.1, seed=2)
    eu = orient.euler_run_dda(6, seed=3)
    m_ps = [np.array([1.5 + 0.01j] * 3), np.array([1.6 + 0.02j] * 3), np.array([1.45 + 0.001j, 1.5 + 0.001j,
                                                                                1.55 + 0.001j])]

    def run():
        out = []
        for m_p in m_ps:
            dd = DiscreteDipoles(Target(name, n, lf, grid, is_in, m_p), IncidentField(0.6, 1.2, eu))
            dd.verbose = False
            dd.tol, dd.itermax = 1e-6, 100
            dd.set_interaction_matrix()
            dd.solve_matrix_equation()
            assert dd.converge
            out.append(dd.X_device.clone())
        return out

    GreenTensor.clear_cache()
    b0 = GreenTensor.builds
    shared = run()
    assert GreenTensor.builds - b0 == 1
    size = GreenTensor.cache_size
    try:
        GreenTensor.cache_size = 0
        GreenTensor.clear_cache()
        fresh = run()
        assert GreenTensor.builds - b0 == 1 + len(m_ps)
    finally:
        GreenTensor.cache_size = size
    for x, y in zip(shared, fresh):
        assert torch.equal(x, y)


def test_sweep_on_the_reference_conditions(tmp_path):
    from bdda_amd import sweep
    from mvp_fft.mvp_fft import GreenTensor
    h5 = _npz("pcas_ocbs_simulated_data.npz")
    ref = _npz("sweep_oracle.npz")
    g = lambda k: h5["target__" + k]  # noqa: E731
    sd = lambda k: h5["target__simulated_data__" + k]  # noqa: E731
    L = sd("C_abs").shape[-1]
    path = str(tmp_path / "sweep.npz")
    store = sweep.create_store(path, g("wl_m_m_pairs"), g("m_p_xyz_list"), g("r_v_base_list"), g("bc_ratio_list"),
                               g("ab_ratio_list"), g("gre_beta_list"), L)
    GreenTensor.clear_cache()
    b0 = GreenTensor.builds
    say("sweep: 16 conditions x 100 orientations")
    summary = sweep.run_sweep(store, log=lambda m: None)
    store.close()
    assert summary["computed"] == 16 and summary["not_converged"] == 0 and set(summary["tries"]) == {1}
    n_shapes, n_pairs = len(g("r_v_base_list")), len(g("wl_m_m_pairs"))
    assert GreenTensor.builds - b0 == n_shapes * n_pairs
    out = sweep.open_store(path)
    o = lambda k: out["target/simulated_data/" + k]  # noqa: E731
    assert out.attrs["target"]["num_orientations"] == L
    assert np.array_equal(o("Euler_angles"), sd("Euler_angles"))
    np.testing.assert_allclose(o("r_ve"), sd("r_ve"), rtol=5e-16)
    # C_abs_mie of a non-absorbing sphere is Q_ext - Q_sca, rounding noise
    # of ~1e-19: compared on the scale of C_ext_mie
    scale = np.max(np.abs(sd("C_ext_mie")))
    for k in ("C_abs_mie", "C_ext_mie"):
        np.testing.assert_allclose(o(k), sd(k), rtol=1e-12, atol=1e-13 * scale)
    for k in ("S_fw_PCAS_mie", "S_bk_OCBS_mie"):
        np.testing.assert_allclose(o(k), sd(k), rtol=1e-6)

    def rel(a, b):  # exact zeros (C_abs of non-absorbing particles) must match exactly
        d = np.abs(a - b)
        return np.max(np.where(b == 0, np.where(d == 0, 0.0, np.inf), d / np.where(b == 0, 1.0, np.abs(b))))

    # per condition, against the oracle: within 1e-6 of the quantity's
    # largest magnitude, or within 10x of how far 1e-13 rad of orientation
    # jitter moves the oracle's own result (ill-conditioned 100 x 100
    # systems on the smallest particles: up to ~3e-4)
    for k in ("C_abs", "C_ext", "S_fw_PCAS_theta", "S_fw_PCAS_phi", "S_bk_OCBS"):
        got, want, sens = o(k), ref[k], ref["sensitivity__" + k]
        for idx in np.ndindex(*want.shape[:-1]):
            den = np.max(np.abs(want[idx]))
            d = np.max(np.abs(got[idx] - want[idx]))
            assert (d == 0 if den == 0 else d / den <= max(1e-6, 10 * sens[idx])), (k, idx, d / den, sens[idx])
        assert rel(got, sd(k)) < 2e-2, k
    # CAS-v2 post-processing of the GPU results classifies every orientation
    # as the oracle's results do
    from bdda_amd import cas
    for ip in range(want.shape[0]):
        for im in range(want.shape[1]):
            got_f = [c["non_sphere_fraction"] for c in cas.cas_v2(out, ip, im)]
            ref_f = [cas.non_sphere_fraction(cas.pcas_depolarization(
                ref["S_fw_PCAS_theta"][ip, im, ir, 0, 0, 0], ref["S_fw_PCAS_phi"][ip, im, ir, 0, 0, 0])
                ["Tot_depol_SNR2"]) for ir in range(want.shape[2])]
            assert got_f == ref_f, (ip, im, got_f, ref_f)
    # resume: a second run finds every condition done
    again = sweep.open_store(path)
    s2 = sweep.run_sweep(again, log=lambda m: None)
    assert s2["computed"] == 0 and s2["skipped"] == 16


def test_retries_and_nan_fill():
    """run_dda's retries (run_dda.py:38-70): a condition whose every try
    fails returns the last try's orientations, NaN observables (real for C,
    complex for S) and the try count; one whose first try fails and second
    converges returns the second try's orientations and results, equal to a
    direct solve at those orientations."""
    from bdda_amd import sweep
    from bl_dda.scatterer import DiscreteDipoles, IncidentField, Target
    name, n, lf, grid, is_in, r_ve = sweep.build_gre_geometry(0.06, 1.2, 1.1, 0.0, sweep.RNG_SEED)
    tg = Target(name, n, lf, grid, is_in, np.array([1.5 + 0.01j] * 3))
    wl, mm, L = 0.4, 1.0, 6
    tries = sweep.orientation_tries(L, sweep.MAX_TRY, sweep.RNG_SEED + 1)
    say("sweep retries: every try fails")
    euler, c_abs, c_ext, s_th, s_ph, s_bk, ok, t = sweep.run_dda(
        tg, wl, mm, L, log=lambda m: None, options={"tol": 1e-12, "itermax": 1, "verbose": False})
    assert not ok and t == sweep.MAX_TRY and np.array_equal(euler, tries[-1])
    assert c_abs.dtype == np.float64 and np.isnan(c_abs).all() and np.isnan(c_ext).all()
    for s in (s_th, s_ph, s_bk):
        assert s.dtype == np.complex128 and np.isnan(s.real).all()
    say("sweep retries: the second try converges")
    first = DiscreteDipoles(tg, IncidentField(wl, mm, tries[0]))
    first.verbose, first.tol, first.itermax = False, 1e-12, 1
    first.set_interaction_matrix()
    first.solve_matrix_equation()
    assert not first.converge
    euler, c_abs, c_ext, s_th, s_ph, s_bk, ok, t = sweep.run_dda(
        tg, wl, mm, L, log=lambda m: None, options={"verbose": False}, first=first)
    assert ok and t == 2 and np.array_equal(euler, tries[1])
    dd = DiscreteDipoles(tg, IncidentField(wl, mm, tries[1]))
    dd.verbose = False
    dd.set_interaction_matrix()
    dd.solve_matrix_equation()
    assert dd.converge
    assert np.array_equal(c_ext, dd.compute_C_ext()) and np.array_equal(s_bk, dd.compute_OCBS_observable_S_bk())
