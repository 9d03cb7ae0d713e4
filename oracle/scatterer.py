"""DDA physics (target, incident field, polarizability, observables) — oracle.

Restates ``bl_dda/scatterer.py`` of the reference in a functional form:

* ``target``           <- ``Target.__init__``                  (``scatterer.py:8-28``)
* ``incident``         <- ``IncidentField.__init__``           (``scatterer.py:31-85``)
* ``polarizability``   <- ``set_interaction_matrix`` :137-148  (CM + CR2009)
* ``rhs``              <- ``set_interaction_matrix`` :153-161
* ``dipoles_from_X``   <- ``solve_matrix_equation`` :181-184
* ``c_abs`` / ``c_ext``<- ``compute_C_abs`` / ``compute_C_ext`` (``:187-199``)
* ``s_fw``             <- ``compute_PCAS_observable_S_fw``     (``:202-213``)
* ``s_bk``             <- ``compute_OCBS_observable_S_bk``     (``:216-228``)
* ``solve``            <- ``DiscreteDipoles`` flow :135-184 with the solver choice

Pinned by: closed forms for a single dipole (P = alpha E_inc, N = 1), the ZYZ
rotation closed form, the optical theorem against Mie in the small-particle
limit, and end-to-end against the reference-held HDF5 fixture
(``dda_results/pcas_ocbs_simulated_data.hdf5``) and ``test_dda.ipynb`` outputs.
"""
import numpy as np
from scipy.spatial.transform import Rotation

from .bl_krylov import bl_bicgstab, bl_cocg_rq
from .mvp_fft import fft_init


class Target:
    """Occupied lattice sites and material (``scatterer.py:8-28``)."""

    def __init__(self, lattice_n, lattice_lf, grid_points, is_in, m_p_xyz):
        self.lattice_n = np.asarray(lattice_n)
        self.lf = lattice_lf
        self.address = np.where(is_in)[0]
        self.pos = grid_points[self.address, :]
        self.N = self.address.size
        self.v = lattice_lf ** 3
        self.m_p_xyz = m_p_xyz
        self.m_p = np.ones_like(self.pos) * m_p_xyz
        self.eps_p = self.m_p ** 2
        self.total_vol = self.v * self.N
        self.ve_radius = np.cbrt(3 * self.total_vol / (4 * np.pi))


class Incident:
    """Lab-frame basis rotated into the particle frame (``scatterer.py:43-85``).

    Rotation ``from_euler('ZYZ', -euler[:, ::-1])`` (``:69-70``); lab vectors
    u_inc = fw = +z, theta = +x, phi = +y, backward u = -z, theta = -x,
    phi = +y (``:51-61``); e0 = theta/sqrt2 + i phi/sqrt2 (``:64-77``).
    """

    def __init__(self, wl_0, m_m, euler):
        self.wl_0, self.m_m = wl_0, m_m
        self.k = 2 * np.pi * m_m / wl_0
        self.L = euler.shape[0]
        R = Rotation.from_euler('ZYZ', euler[:, ::-1] * -1.0).as_matrix()
        ex, ey, ez = np.eye(3)
        self.rot = R
        self.u_inc = R @ ez
        self.th_inc = R @ ex
        self.ph_inc = R @ ey
        self.e0 = (1 / np.sqrt(2)) * self.th_inc + (1j / np.sqrt(2)) * self.ph_inc
        self.u_fw, self.th_fw, self.ph_fw = R @ ez, R @ ex, R @ ey
        self.u_bk, self.th_bk, self.ph_bk = R @ (-ez), R @ (-ex), R @ ey


def polarizability(tg, inc):
    """CM + CR2009 polarizability and diag(A) = 1/alpha (``scatterer.py:137-148``).

    Returns (eps_r (N,3), alpha (N,3), diag_A (3N,)).
    """
    eps_r = tg.eps_p / inc.m_m ** 2
    a0 = (3 / (4 * np.pi)) * ((eps_r - 1) / (eps_r + 2)) * tg.v
    a = (3 * tg.v / (4 * np.pi)) ** (1 / 3)
    Mt = (8 * np.pi / 3) * ((1 - 1j * inc.k * a) * np.exp(1j * inc.k * a) - 1)
    alpha = a0 / (1 - Mt * a0 / tg.v)
    return eps_r, alpha, (1.0 / alpha).ravel()


def rhs(tg, inc):
    """E_inc (L,N,3) and B (3N, L), ``scatterer.py:155-161``."""
    phase = np.exp(1j * inc.k * (inc.u_inc @ tg.pos.T))
    E_inc = phase[:, :, None] * inc.e0[:, None, :]
    return E_inc, E_inc.reshape(inc.L, 3 * tg.N).T


def dipoles_from_X(tg, inc, X, eps_r):
    """P (L,N,3) and E (L,N,3) from X (3N,L), ``scatterer.py:182-184``."""
    P = X.T.reshape(inc.L, tg.N, 3)
    E = P * (4 * np.pi) / ((eps_r[None, :, :] - 1) * tg.v)
    return P, E


def c_abs(inc, P, E, eps_r):
    """``scatterer.py:187-193`` (zeros unless mean Im eps_r >= 1e-12)."""
    if np.mean(eps_r.imag) >= 1e-12:
        return 4 * np.pi * inc.k * np.sum(np.imag(P * np.conj(E)), axis=(1, 2))
    return np.zeros(inc.L)


def c_ext(inc, P, E_inc):
    """``scatterer.py:196-199``."""
    return 4 * np.pi * inc.k * np.sum(np.imag(P * np.conj(E_inc)), axis=(1, 2))


def _amp(inc, tg, P, u, th, ph):
    e = np.exp(1j * inc.k * (u @ tg.pos.T))
    pt = np.einsum('lni,li->ln', P, th)
    pp = np.einsum('lni,li->ln', P, ph)
    s2 = np.sqrt(2)
    return (inc.k ** 2 * (pt * np.conj(s2 * e)).sum(axis=1),
            inc.k ** 2 * (pp * np.conj(s2 * 1j * e)).sum(axis=1))


def s_fw(tg, inc, P):
    """(S_theta, S_phi) forward, ``scatterer.py:202-213``."""
    return _amp(inc, tg, P, inc.u_fw, inc.th_fw, inc.ph_fw)


def s_bk(tg, inc, P):
    """S_bk = (-S_theta + S_phi)/sqrt2 backward, ``scatterer.py:216-228``."""
    st, sp = _amp(inc, tg, P, inc.u_bk, inc.th_bk, inc.ph_bk)
    return (-st + sp) / np.sqrt(2)


def solve(tg, inc, solver="bicgstab", tol=1e-2, itermax=25, X0=None):
    """Full DDA solve; returns a dict with X, iter_fin, err_fin, history and
    the observables (C_abs, C_ext, S_fw_theta, S_fw_phi, S_bk, converge)."""
    eps_r, alpha, diag = polarizability(tg, inc)
    At = fft_init(tg.lattice_n, 3, tg.lf, inc.k)
    E_inc, B = rhs(tg, inc)
    fn = bl_bicgstab if solver == "bicgstab" else bl_cocg_rq
    X, it, err, hist = fn(tg.lattice_n, 3, tg.address, At, diag, B, tol, itermax, X0=X0)
    out = dict(X=X, iter_fin=it, err_fin=err, history=hist, diag=diag, B=B,
               A_tilde=At, converge=err < tol)
    P, E = dipoles_from_X(tg, inc, X, eps_r)
    out.update(P=P, E=E, C_abs=c_abs(inc, P, E, eps_r), C_ext=c_ext(inc, P, E_inc))
    out["S_fw_theta"], out["S_fw_phi"] = s_fw(tg, inc, P)
    out["S_bk"] = s_bk(tg, inc, P)
    return out
