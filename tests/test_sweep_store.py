"""The sweep driver's results file on the CPU: the layout that
``create_store`` writes equals the reference's stored file
(``dda_results/pcas_ocbs_simulated_data.hdf5``, created by
``dda_results/create_h5py.ipynb``): same dataset paths, shapes and dtypes;
an .npz store survives a write / reopen round trip; the orientation draws
per try follow ``run_dda.py:46-54``."""
import numpy as np

from bdda_amd import sweep
from oracle import orient


def test_layout_matches_the_reference_file(golden_h5, tmp_path):
    h5 = golden_h5
    g = lambda k: h5["target__" + k]  # noqa: E731
    L = h5["target__simulated_data__C_abs"].shape[-1]
    st = sweep.create_store(str(tmp_path / "r.npz"), g("wl_m_m_pairs"), g("m_p_xyz_list"), g("r_v_base_list"),
                            g("bc_ratio_list"), g("ab_ratio_list"), g("gre_beta_list"), L)
    ours = {k.replace("/", "__"): v for k, v in st.arrays.items()}
    assert set(ours) == set(h5)
    for k, v in h5.items():
        assert ours[k].shape == v.shape and ours[k].dtype == v.dtype, k
        if not k.startswith("target__simulated_data"):
            assert np.array_equal(ours[k], v), k
    assert st.attrs["target"]["num_orientations"] == L
    assert "definition" in st.attrs["target/simulated_data/S_bk_OCBS"]


def test_npz_store_round_trip(tmp_path):
    path = str(tmp_path / "s.npz")
    st = sweep.create_store(path, [[0.6, 1.33]], [[1.5 + 0.01j] * 3], [0.1, 0.2], [1.0], [1.0], [0.0], 7)
    st["target/simulated_data/S_fw_PCAS_mie"][0, 0, 1, 0, 0, 0] = 1 + 2j
    st["target/simulated_data/Euler_angles"][0, 0, 1, 0, 0, 0] = np.arange(21).reshape(7, 3)
    st.flush()
    back = sweep.open_store(path)
    assert back["target/simulated_data/S_fw_PCAS_mie"][0, 0, 1, 0, 0, 0] == 1 + 2j
    assert np.array_equal(back["target/simulated_data/Euler_angles"][0, 0, 1, 0, 0, 0],
                          np.arange(21).reshape(7, 3))
    assert back.attrs == st.attrs


def test_orientation_tries_follow_the_reference_draws(golden_h5):
    tries = sweep.orientation_tries(100, 4, 12346)
    for t in range(4):
        assert np.array_equal(tries[t], orient.euler_run_dda(100, seed=12346, tries=t + 1))
    stored = golden_h5["target__simulated_data__Euler_angles"]
    assert np.array_equal(tries[0], stored[0, 0, 0, 0, 0, 0])
