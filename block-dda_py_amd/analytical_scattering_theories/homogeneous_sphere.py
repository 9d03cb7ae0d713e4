"""Mie theory for a homogeneous sphere (host-side physics check).

Drop-in for the reference's ``analytical_scattering_theories/
homogeneous_sphere.py`` (``mie_compute_q_and_s``, lines 4-117), which
``test_dda.ipynb`` (cell 0, 9) and ``run_dda.py:7,124-128`` import to compare
DDA results with the volume-equivalent sphere.  Same arguments, conventions
and return values: efficiencies Q_sca, Q_abs, Q_ext (float) and the PCAS
forward amplitude S_fw = (S11 + S22)(0)/2 and OCBS backscatter amplitude
S_bk = (-S11 + S22)(pi)/sqrt 2 cast to complex64, with S11 = S2/(-ik),
S22 = S1/(-ik) (Bohren & Huffman 1983, chapter 4).

Formulation: the Riccati-Bessel functions of the size parameter come from
scipy's spherical Bessel functions (psi_n = x j_n(x), xi_n = x h1_n(x)); the
logarithmic derivative D_n(mx) from a downward continued recurrence started
15 orders beyond max(n_stop, |mx|) (stable for absorbing spheres); the
angular functions pi_n, tau_n from their three-term recurrences.  The series
stops at n_stop = floor(x + 4 x^0.3333 + 2), the reference's (and BH83's)
truncation, so results agree with the reference to float64 rounding.
"""
import numpy as np
from scipy.special import spherical_jn, spherical_yn


def _n_stop(x):
    return int(np.floor(abs(x + 4.0 * x ** 0.3333 + 2.0)))


def _log_derivative(mx, n_stop):
    """D_n(mx) = psi_n'(mx) / psi_n(mx) for n = 0 .. n_stop, by the downward
    recurrence D_{n-1} = n/z - 1/(D_n + n/z) from D_start = 0."""
    start = int(np.floor(max(n_stop, abs(mx)) + 15))
    d = 0.0 + 0.0j
    out = np.empty(n_stop + 1, dtype=np.complex128)
    for n in range(start, 0, -1):
        d = n / mx - 1.0 / (d + n / mx)
        if n - 1 <= n_stop:
            out[n - 1] = d
    return out


def _riccati_bessel(x, n_stop):
    """psi_n(x) and xi_n(x) = psi_n(x) + i x y_n(x), n = 0 .. n_stop."""
    n = np.arange(n_stop + 1)
    psi = x * spherical_jn(n, x)
    return psi, psi + 1j * (x * spherical_yn(n, x))


def _angular(mu, n_stop):
    """pi_n(mu), tau_n(mu) for n = 1 .. n_stop, shape (n_stop, len(mu))."""
    pi_n = np.zeros((n_stop + 1, mu.size))
    tau_n = np.zeros_like(pi_n)
    prev, cur = np.zeros_like(mu), np.ones_like(mu)  # pi_0, pi_1
    for n in range(1, n_stop + 1):
        pi_n[n] = cur
        tau_n[n] = n * mu * cur - (n + 1) * prev
        prev, cur = cur, ((2 * n + 1) * mu * cur - (n + 1) * prev) / n
    return pi_n[1:], tau_n[1:]


def mie_coefficients(x, m_r):
    """a_n, b_n (n = 1 .. n_stop) for size parameter x and relative index m_r."""
    n_stop = _n_stop(x)
    D = _log_derivative(m_r * x, n_stop)[1:]
    psi, xi = _riccati_bessel(x, n_stop)
    n = np.arange(1, n_stop + 1, dtype=float)
    ta = D / m_r + n / x
    tb = m_r * D + n / x
    a = (ta * psi[1:] - psi[:-1]) / (ta * xi[1:] - xi[:-1])
    b = (tb * psi[1:] - psi[:-1]) / (tb * xi[1:] - xi[:-1])
    return a, b


def mie_compute_q_and_s(wl_0, m_m, r_p, m_p, nang=3):
    """(Q_sca, Q_abs, Q_ext, S_fw complex64, S_bk complex64) of a sphere of
    radius r_p and index m_p in a medium of index m_m at vacuum wavelength
    wl_0; `nang` scattering angles from 0 to 180 degrees (homogeneous_sphere.py:4-117)."""
    nang = int(nang)
    k = m_m * 2.0 * np.pi / wl_0
    x = k * r_p
    a, b = mie_coefficients(x, m_p / m_m)
    n = np.arange(1, a.size + 1, dtype=float)
    w = 2.0 * n + 1.0
    q_sca = float(np.real(2.0 / x ** 2 * np.sum(w * (np.abs(a) ** 2 + np.abs(b) ** 2))))
    q_ext = float(np.real(2.0 / x ** 2 * np.sum(w * (a + b))))
    q_abs = q_ext - q_sca
    mu = np.cos(np.linspace(0, np.pi, nang))
    pi_n, tau_n = _angular(mu, a.size)
    c = w / (n * (n + 1.0))
    S1 = (c * a) @ pi_n + (c * b) @ tau_n
    S2 = (c * a) @ tau_n + (c * b) @ pi_n
    S11, S22 = S2 / (-1j * k), S1 / (-1j * k)
    s_fw = np.complex64((S11[0] + S22[0]) / 2.0)
    s_bk = np.complex64((-S11[-1] + S22[-1]) / np.sqrt(2))
    return q_sca, q_abs, q_ext, s_fw, s_bk
