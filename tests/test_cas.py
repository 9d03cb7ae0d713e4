"""CAS-v2 post-processing (bdda_amd.cas) against the reference's recorded
outputs: plot_dda_results.ipynb cell 8 prints the non-sphere fractions
0, 0, 0.57, 0.53 of the four particle sizes of one condition of the stored
sweep, and cell 4 their r_ve 0.0999, 0.2027, 0.3065, 0.4095 um.  The
fractions are those of condition (pair 1, m_p 1) of the stored file, all
four exactly (the notebook's cell 4 as committed selects (0, 0), so it was
last run with another selection; (0, 0) gives 0, 0, 0.46, 0.49).  Plus
closed-form checks of the depolarization formulas."""
import numpy as np

from bdda_amd import cas, sweep


class _Store:
    """The golden npz (datasets keyed target__...) behind store paths."""

    def __init__(self, d):
        self.d = d

    def __getitem__(self, key):
        return self.d[key.replace("/", "__")]


def test_notebook_non_sphere_fractions(golden_h5):
    out = cas.cas_v2(_Store(golden_h5), i_pair=1, i_mp=1)
    assert [round(o["non_sphere_fraction"], 4) for o in out] == [0.0, 0.0, 0.57, 0.53]
    others = [[round(o["non_sphere_fraction"], 4) for o in cas.cas_v2(_Store(golden_h5), ip, im)]
              for ip, im in ((0, 0), (0, 1), (1, 0))]
    assert [0.0, 0.0, 0.57, 0.53] not in others
    assert [round(o["r_ve"], 4) for o in out] == [0.0999, 0.2027, 0.3065, 0.4095]


def test_depolarization_closed_forms():
    # an isotropic response (s = p) has zero depolarization
    d = cas.pcas_depolarization(np.array([0.1 + 0.2j]), np.array([0.1 + 0.2j]))
    assert d["ReS_depol"][0] == 0 and d["Tot_depol_SNR2"][0] == 0
    # s = 3a, p = a: depol 1/2, mean 2a, rms sqrt(1.25/2) * noise / (2a)
    a = 0.05
    d = cas.pcas_depolarization(np.array([3 * a + 3j * a]), np.array([a + 1j * a]))
    np.testing.assert_allclose(d["ReS_depol"], 0.5)
    snr_re = 0.5 / (np.sqrt(1.25 / 2) * cas.RES_RMS / (2 * a))
    snr_im = 0.5 / (np.sqrt(1.25 / 2) * cas.IMS_RMS / (2 * a))
    np.testing.assert_allclose(d["Tot_depol_SNR2"], snr_re ** 2 + snr_im ** 2)
    assert cas.non_sphere_fraction([1.0, 20.0, 13.82, 14.0]) == 0.5


def test_on_a_sweep_store(tmp_path, golden_h5):
    """The same summary read from a results file written by the sweep driver."""
    g = lambda k: golden_h5["target__" + k]  # noqa: E731
    st = sweep.create_store(str(tmp_path / "r.npz"), g("wl_m_m_pairs"), g("m_p_xyz_list"), g("r_v_base_list"),
                            g("bc_ratio_list"), g("ab_ratio_list"), g("gre_beta_list"), 100)
    for k in ("S_fw_PCAS_theta", "S_fw_PCAS_phi", "r_ve", "S_fw_PCAS_mie", "Euler_angles"):
        st.arrays["target/simulated_data/" + k][...] = golden_h5["target__simulated_data__" + k]
    st.flush()
    out = cas.cas_v2(sweep.open_store(st.path), i_pair=1, i_mp=1)
    assert [round(o["non_sphere_fraction"], 4) for o in out] == [0.0, 0.0, 0.57, 0.53]
