"""The drop-in Mie module (``analytical_scattering_theories.homogeneous_sphere``)
against the reference's recorded notebook values, closed forms and the
float64 oracle.

``test_dda.ipynb`` cell 9 recorded Q_sca, Q_abs, Q_ext at 16 digits and the
complex64 S_fw, S_bk of the volume-equivalent sphere (r_ve of the 17^3
notebook lattice, m_p = 1.5 + 0.01i, m_m = 1.329, wl_0 = 0.834 um).
"""
import numpy as np
import pytest

from analytical_scattering_theories.homogeneous_sphere import mie_compute_q_and_s
from oracle import gre, mie, scatterer


def test_notebook_values(golden_nb):
    nb = golden_nb
    _, n, lf, grid, is_in, _, _ = gre.build_gre_geometry(float(nb["r_v_base"]), 1, 1, 0.0,
                                                         seed=int(nb["shape_seed"]))
    r_ve = scatterer.Target(n, lf, grid, is_in, nb["m_p_xyz"]).ve_radius
    q = mie_compute_q_and_s(float(nb["wl_0"]), float(nb["m_m"]), r_ve, 1.5 + 0.01j, nang=3)
    np.testing.assert_allclose(q[:3], nb["mie_q"], rtol=1e-13)
    assert q[3] == nb["mie_s_fw"] and q[4] == nb["mie_s_bk"]
    assert q[3].dtype == np.complex64 and q[4].dtype == np.complex64


@pytest.mark.parametrize("args", [(0.834, 1.329, 0.2, 1.5 + 0.01j), (0.5, 1.0, 1.3, 1.6 + 0.1j),
                                  (0.55, 1.33, 0.02, 1.5 + 1e-4j), (0.5, 1.0, 3.0, 1.33 + 0.5j),
                                  (0.6, 1.0, 0.8, 2.5 + 1.0j)])
@pytest.mark.parametrize("nang", [3, 7])
def test_against_oracle(args, nang):
    a = mie_compute_q_and_s(*args, nang=nang)
    b = mie.mie_q_and_s(*args, nang=nang)
    np.testing.assert_allclose(a[:3], b[:3], rtol=1e-12)
    for u, v in zip(a[3:], b[3:]):
        assert abs(u - v) <= 2e-7 * abs(v)


def test_rayleigh_limit_and_optical_theorem():
    wl, mm, r, mp = 10.0, 1.0, 0.05, 1.5 + 0.1j
    Qs, Qa, Qe, Sf, Sb = mie_compute_q_and_s(wl, mm, r, mp)
    x = 2 * np.pi * r / wl
    K = (mp ** 2 - 1) / (mp ** 2 + 2)
    np.testing.assert_allclose(Qa, 4 * x * K.imag, rtol=2e-3)
    np.testing.assert_allclose(Qs, 8 / 3 * x ** 4 * abs(K) ** 2, rtol=2e-2)
    # C_ext = (4 pi / k) Im S(0) with S11 = S2 / (-ik)
    wl, mm, r = 0.6, 1.33, 0.3
    Qs, Qa, Qe, Sf, Sb = mie_compute_q_and_s(wl, mm, r, 1.5 + 0.01j)
    k = 2 * np.pi * mm / wl
    np.testing.assert_allclose(Qe * np.pi * r ** 2, 4 * np.pi / k * float(np.imag(Sf)), rtol=1e-6)
    # a non-absorbing sphere absorbs nothing
    Qs, Qa, Qe, _, _ = mie_compute_q_and_s(0.6, 1.0, 0.4, 1.5 + 0.0j)
    assert abs(Qa) < 1e-12 * Qe
