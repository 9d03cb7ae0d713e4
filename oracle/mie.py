"""Mie theory for a homogeneous sphere — oracle (physics check only).

Restates ``analytical_scattering_theories/homogeneous_sphere.py:4-117``
(Bohren & Huffman 1983): D_n downward recurrence (``:46-49``), psi_n via the
ratio recurrence (``:52-61``), chi_n upward (``:64-68``), a_n, b_n (``:79-82``),
efficiencies (``:85-88``), pi_n/tau_n (``:91-102``) and the PCAS/OCBS
amplitudes cast to complex64 (``:105-115``).

Pinned by: the ``test_dda.ipynb`` recorded Q_sca, Q_abs, Q_ext, S_fw, S_bk at
full precision (``test_dda.ipynb`` cell 9), the Rayleigh limit and the optical
theorem Q_ext = 4 Re S(0) / x^2 * ... (tests).
"""
import numpy as np


def mie_q_and_s(wl_0, m_m, r_p, m_p, nang=3):
    """Returns (Qsca, Qabs, Qext, S_fw complex64, S_bk complex64)."""
    nang = int(nang)
    k = m_m * 2.0 * np.pi / wl_0
    x = k * r_p
    mr = m_p / m_m
    nstop = int(np.floor(abs(x + 4.0 * x ** 0.3333 + 2.0)))
    y = mr * x
    nmx = int(np.floor(max(nstop, abs(y)) + 15))
    D = np.zeros(nmx + 1, dtype=np.complex128)
    for n in range(nmx, 0, -1):
        D[n - 1] = n / y - 1.0 / (D[n] + n / y)
    D = D[:nstop + 1]
    Rt = np.zeros(nmx + 1)
    Rt[nmx] = x / (2 * nmx + 1)
    for n in range(nmx - 1, -1, -1):
        Rt[n] = 1.0 / ((2 * n + 1) / x - Rt[n + 1])
    psi = np.empty(nstop + 1)
    psi[0] = Rt[0] * np.cos(x)
    psi[1:] = psi[0] * np.cumprod(Rt[1:nstop + 1])
    chi = np.zeros(nstop + 1)
    chi[0] = -np.cos(x)
    chi[1] = chi[0] / x - np.sin(x)
    for n in range(2, nstop + 1):
        chi[n] = ((2 * n - 1) / x) * chi[n - 1] - chi[n - 2]
    xi = psi + 1j * chi
    nn = np.arange(1, nstop + 1, dtype=float)
    Dn = D[1:]
    nx = nn / x
    an = ((Dn / mr + nx) * psi[1:] - psi[:-1]) / ((Dn / mr + nx) * xi[1:] - xi[:-1])
    bn = ((mr * Dn + nx) * psi[1:] - psi[:-1]) / ((mr * Dn + nx) * xi[1:] - xi[:-1])
    f1 = 2.0 * nn + 1.0
    Qsca = float(np.real((2.0 / x ** 2) * np.sum(f1 * (np.abs(an) ** 2 + np.abs(bn) ** 2))))
    Qext = float(np.real((2.0 / x ** 2) * np.sum(f1 * (an + bn))))
    Qabs = Qext - Qsca
    mu = np.cos(np.linspace(0, np.pi, nang))
    pie = np.zeros((nstop + 1, nang))
    tau = np.zeros((nstop + 1, nang))
    pie[1] = 1.0
    pie[2] = 3.0 * mu
    tau[1] = mu
    tau[2] = 6.0 * mu ** 2 - 3.0
    for n in range(3, nstop + 1):
        pie[n] = ((2 * n - 1) / (n - 1)) * mu * pie[n - 1] - (n / (n - 1)) * pie[n - 2]
        tau[n] = n * mu * pie[n] - (n + 1) * pie[n - 1]
    f2 = (2.0 * nn + 1.0) / (nn * (nn + 1.0))
    S1 = (f2 * an) @ pie[1:] + (f2 * bn) @ tau[1:]
    S2 = (f2 * an) @ tau[1:] + (f2 * bn) @ pie[1:]
    S11 = S2 / (-1j * k)
    S22 = S1 / (-1j * k)
    return (Qsca, Qabs, Qext, np.complex64((S11[0] + S22[0]) / 2.0),
            np.complex64((-S11[-1] + S22[-1]) / np.sqrt(2)))
