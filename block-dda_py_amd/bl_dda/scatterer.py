"""DDA physics API on the GPU — drop-in for the reference ``bl_dda/scatterer.py``.

``Target``, ``IncidentField`` and ``DiscreteDipoles`` keep the reference's
constructor arguments, attributes, methods and return tuples
(``scatterer.py:8-228``).  The host keeps the O(N) and O(L) set-up the
reference does in NumPy (occupied addresses, polarizability, rotations);
the Green tensor FFT, the right-hand side, the block-Krylov solve and the
observable reductions run in libbdda's HIP kernels, and the Krylov blocks
stay on the device (``X``, ``B``, ``P``, ``E``, ``E_inc`` are fetched to the
host only when read).

Extensions (attributes, set before ``solve_matrix_equation``):
``solver`` ('bicgstab' as the reference, or 'cocg_rq'), ``precision``
('fp64' or 'mixed'), ``refine_every`` (20), ``X0`` (warm start, (3N, L)),
``pad_to`` (compute on the lattice padded with vacuum cells to at least
this many cells per axis; without it the solver's cuboid is the smallest box
of lattice cells holding the occupied sites, ``solver_dims`` after
``set_interaction_matrix``; the physics is unchanged either way),
``verbose``.
"""
import time

import numpy as np
from scipy.spatial.transform import Rotation as R

from bdda_amd.plan import Plan
from mvp_fft.mvp_fft import GreenTensor


class Target:
    """Occupied sites and homogeneous (optionally anisotropic) material (scatterer.py:8-28)."""

    def __init__(self, shape_name, lattice_n, lattice_lf, lattice_grid_points, lattice_grid_points_is_in_target,
                 m_p_xyz):
        self.shape_name = shape_name
        self.lattice_n = lattice_n
        self.lattice_lf = lattice_lf
        self.lattice_grid_points = lattice_grid_points
        self.lattice_grid_points_is_in_target = lattice_grid_points_is_in_target
        self.m_p_xyz = m_p_xyz
        self.lattice_address_in_target = np.where(lattice_grid_points_is_in_target)[0]
        self.lattice_pos_in_target = lattice_grid_points[self.lattice_address_in_target, :]
        self.num_element_occupy = self.lattice_address_in_target.size
        self.element_vol = self.lattice_lf ** 3
        self.m_p = np.ones_like(self.lattice_pos_in_target) * m_p_xyz
        self.eper_p = self.m_p ** 2
        self.total_vol = self.element_vol * self.num_element_occupy
        self.ve_radius = np.cbrt(3 * self.total_vol / (4 * np.pi))


class IncidentField:
    """Lab-frame incident/scattering bases rotated into the particle frame (scatterer.py:31-85)."""

    def __init__(self, wl_0, m_m, euler_angles):
        self.wl_0 = wl_0
        self.m_m = m_m
        self.eper_m = m_m ** 2
        self.k = 2 * np.pi * m_m / wl_0
        self.euler_angles = euler_angles
        self.L = euler_angles.shape[0]
        ex = np.array([1.0, 0.0, 0.0])
        ey = np.array([0.0, 1.0, 0.0])
        ez = np.array([0.0, 0.0, 1.0])
        self.e0_inc_pol_theta = 1 / np.sqrt(2)
        self.e0_inc_pol_phi = 1j / np.sqrt(2)
        # lab -> particle: inverse of the particle orientation, ZYZ(-g, -b, -a)
        rot = R.from_euler('ZYZ', euler_angles[:, ::-1] * -1.0).as_matrix()
        self.u_inc_vec_P = rot @ ez
        self.theta_inc_vec_P = rot @ ex
        self.phi_inc_vec_P = rot @ ey
        self.e0_inc_vec_P = self.e0_inc_pol_theta * self.theta_inc_vec_P + self.e0_inc_pol_phi * self.phi_inc_vec_P
        self.u_sca_fw_vec_P = rot @ ez
        self.theta_sca_fw_vec_P = rot @ ex
        self.phi_sca_fw_vec_P = rot @ ey
        self.u_sca_bk_vec_P = rot @ (-ez)
        self.theta_sca_bk_vec_P = rot @ (-ex)
        self.phi_sca_bk_vec_P = rot @ ey


class DiscreteDipoles(Target, IncidentField):
    """DDA system on the GPU (scatterer.py:88-228).

    ``DiscreteDipoles.verbose_default`` (True, like the reference's
    unconditional prints) sets the initial ``verbose`` of new instances.
    """

    verbose_default = True

    def __init__(self, target, incidentfield):
        Target.__init__(self, target.shape_name, target.lattice_n, target.lattice_lf, target.lattice_grid_points,
                        target.lattice_grid_points_is_in_target, target.m_p_xyz)
        IncidentField.__init__(self, incidentfield.wl_0, incidentfield.m_m, incidentfield.euler_angles)
        N = self.num_element_occupy
        self.dpl = (self.wl_0 / np.abs(np.max(self.m_p_xyz))) / self.lattice_lf
        self.f = 3
        self.eper_r = np.zeros((N, 3), dtype=np.complex128)
        self.alpha_E = np.zeros((N, 3), dtype=np.complex128)
        self.diag_A = np.zeros(3 * N, dtype=np.complex128)
        self.itermax = 25
        self.tol = 1e-2
        self.converge = False
        self.C_abs = None
        self.C_ext = None
        self.S_fw_PCAS_theta = None
        self.S_fw_PCAS_phi = None
        self.S_bk_OCBS_theta = None
        self.S_bk_OCBS_phi = None
        self.S_bk_OCBS = None
        self.S_PCAS_sp_avg = None
        self.S_PCAS_sp_avg_SNR = None
        self.S_PCAS_depol = None
        self.S_PCAS_depol_SNR = None
        # extensions
        self.solver = "bicgstab"
        self.precision = "fp64"
        self.refine_every = 20
        # a float64 solve stops on its recursive residual, as the reference's
        # does; True recomputes the true residual within the solve (one more
        # operator application) instead of on first access
        self.check_true_residual = False
        self.X0 = None
        self.pad_to = None
        self.solver_dims = None
        self.verbose = DiscreteDipoles.verbose_default
        self.iter_fin = None
        self.err_fin = None
        self._true_residual = None
        self.err_history = []
        self.warm_start_used = False
        self.solve_time = None
        self._X_dev = None
        self._B_dev = None
        self._obs = None
        self.Au_til = None
        if self.verbose:
            print("Number of dipoles per wavelength in the particle volume: dpl= {:}".format(self.dpl))

    # ---------------------------------------------------------- set-up
    def _compute_dims(self, ix, iy, iz):
        """Solver cuboid (cells per axis) and the lattice cell at its origin.

        With ``pad_to`` the lattice itself, padded with vacuum cells to at
        least ``pad_to``.  Without it the smallest box of lattice cells that
        holds every occupied site: the enclosing lattice keeps a margin of
        vacuum cells round the particle (up to a third of an axis), which
        would only enlarge the FFT grid.  Either way the convolution over
        the occupied sites is exact and the physics is unchanged."""
        n = np.array([int(v) for v in self.lattice_n])
        if self.pad_to is not None:
            return np.maximum(n, np.asarray(self.pad_to, dtype=int)), np.zeros(3, dtype=np.int64)
        if ix.size == 0:
            return n, np.zeros(3, dtype=np.int64)
        lo = np.array([ix.min(), iy.min(), iz.min()], dtype=np.int64)
        hi = np.array([ix.max(), iy.max(), iz.max()], dtype=np.int64) + 1
        return hi - lo, lo

    def set_interaction_matrix(self):
        """Polarizability, diag(A), device Green spectrum and B (scatterer.py:135-161)."""
        self.eper_r = self.eper_p / self.eper_m
        alpha0 = (3 / (4 * np.pi)) * ((self.eper_r - 1) / (self.eper_r + 2)) * self.element_vol
        a = (3 * self.element_vol / (4 * np.pi)) ** (1 / 3)
        M_term = (8 * np.pi / 3) * ((1 - 1j * self.k * a) * np.exp(1j * self.k * a) - 1)
        self.alpha_E = alpha0 / (1 - M_term * alpha0 / self.element_vol)
        self.diag_A = (1.0 / self.alpha_E).ravel()
        self.jpre = 1.0 / self.diag_A
        nx0, ny0, nz0 = (int(v) for v in self.lattice_n)
        addr = self.lattice_address_in_target.astype(np.int64)
        ix, rem = np.divmod(addr, ny0 * nz0)
        iy, iz = np.divmod(rem, nz0)
        dims, lo = self._compute_dims(ix, iy, iz)
        if tuple(dims) != (nx0, ny0, nz0):
            addr = ((ix - lo[0]) * dims[1] + (iy - lo[1])) * dims[2] + (iz - lo[2])
        self._plan_address = addr
        self.solver_dims = tuple(int(v) for v in dims)
        plan = Plan(dims)
        self.Au_til = GreenTensor(dims, self.lattice_lf, self.k, plan=plan)
        plan.set_target(addr, self.diag_A, self.jpre, pos=self.lattice_pos_in_target,
                        key=("dd", id(self)))
        self._plan = plan
        self._B_dev = plan.rhs(self.u_inc_vec_P, self.e0_inc_vec_P, self.k)
        self._X_dev = None
        self._obs = None

    @property
    def B(self):
        return None if self._B_dev is None else self._B_dev.cpu().numpy()

    @property
    def e_inc_phase(self):
        return np.exp(1j * self.k * (self.u_inc_vec_P @ self.lattice_pos_in_target.T))

    @property
    def E_inc(self):
        return self.B.T.reshape(self.L, self.num_element_occupy, 3)

    # ----------------------------------------------------------- solve
    def solve_matrix_equation(self):
        """Device block-Krylov solve (scatterer.py:164-184)."""
        start = time.time()
        if self.verbose:
            print("Starting {} iterative solver...".format(self._solver_name()))
        X, it, err, tr, hist, warm = self._plan.solve(self._B_dev, self.solver, self.precision, self.tol,
                                                      self.itermax, X0=self.X0, refine_every=self.refine_every,
                                                      true_residual=self.check_true_residual)
        self.solve_info = dict(self._plan.last_solve_info)
        self._store_solution(X, it, err, tr, hist, warm, time.time() - start)

    @property
    def true_residual(self):
        """||jpre (B - A X)|| / ||jpre B|| of the current solution: a mixed
        solve reports it, a float64 one (which, like the reference, stops on
        its recursive residual) computes it on first access."""
        if self._true_residual is None and self._X_dev is not None:
            self._true_residual = self._plan.true_residual(self._B_dev, self._X_dev)
        return self._true_residual

    @true_residual.setter
    def true_residual(self, v):
        self._true_residual = v

    def _solver_name(self):
        return "block-BiCGStab" if self.solver == "bicgstab" else "block-COCG-rq"

    def _store_solution(self, X, it, err, tr, hist, warm, seconds):
        """Solver outputs -> attributes, reporting as scatterer.py:173-184."""
        if not hist:
            err = float('inf')
        self.solve_time = seconds
        self.iter_fin, self.err_fin, self.true_residual = it, err, tr
        self.err_history, self.warm_start_used = hist, warm
        if self.verbose:
            for k, e in enumerate(hist):
                print("iter= {:}, err= {:.4f}".format(k, e))
        self._X_dev = X
        self._obs = None
        self.converge = err < self.tol
        if self.converge and self.verbose:
            print("{} converged! (iter_fin={:}, err_fin={:.4f}, solver time={:.1f}s)".format(
                self._solver_name(), it, err, self.solve_time))

    @property
    def X(self):
        if self._X_dev is None:
            return np.zeros((3 * self.num_element_occupy, self.L), dtype=np.complex128)
        return self._X_dev.cpu().numpy()

    @property
    def X_device(self):
        return self._X_dev

    @property
    def P(self):
        if not self.converge:
            return np.zeros((self.L, self.num_element_occupy, 3), dtype=np.complex128)
        return self.X.T.reshape(self.L, self.num_element_occupy, 3)

    @property
    def E(self):
        if not self.converge:
            return np.zeros((self.L, self.num_element_occupy, 3), dtype=np.complex128)
        return self.P * (4 * np.pi) / ((self.eper_r[np.newaxis, :, :] - 1) * self.element_vol)

    # ------------------------------------------------------ observables
    def _observable_sums(self):
        if self._obs is None:
            if not self.converge or self._X_dev is None:
                self._obs = np.zeros((self.L, 5), dtype=np.complex128)
            else:
                e0 = self.e0_inc_vec_P
                geo = np.concatenate([self.u_sca_fw_vec_P, self.theta_sca_fw_vec_P, self.phi_sca_fw_vec_P,
                                      self.u_sca_bk_vec_P, self.theta_sca_bk_vec_P, self.phi_sca_bk_vec_P,
                                      np.stack([e0.real, e0.imag], axis=-1).reshape(self.L, 6)], axis=1)
                # E = P * 4 pi / ((eps_r - 1) v) per axis (homogeneous material)
                epsfac = (4 * np.pi) / ((self.eper_r[0] - 1) * self.element_vol)
                self._obs = self._plan.observables(self._X_dev, geo, epsfac, self.k)
        return self._obs

    def compute_C_abs(self):
        """scatterer.py:187-193."""
        self.C_abs = np.zeros(self.L, dtype=np.float64)
        if np.mean(self.eper_r.imag) >= 1e-12:
            self.C_abs = 4 * np.pi * self.k * self._observable_sums()[:, 4].real
        return self.C_abs

    def compute_C_ext(self):
        """scatterer.py:196-199."""
        self.C_ext = 4 * np.pi * self.k * self._observable_sums()[:, 4].imag
        return self.C_ext

    def compute_PCAS_observable_S_fw(self):
        """scatterer.py:202-213."""
        o = self._observable_sums()
        s2 = np.sqrt(2)
        self.S_fw_PCAS_theta = self.k ** 2 * s2 * o[:, 0]
        self.S_fw_PCAS_phi = self.k ** 2 * s2 * o[:, 1]
        return self.S_fw_PCAS_theta, self.S_fw_PCAS_phi

    def compute_OCBS_observable_S_bk(self):
        """scatterer.py:216-228."""
        o = self._observable_sums()
        s2 = np.sqrt(2)
        self.S_bk_OCBS_theta = self.k ** 2 * s2 * o[:, 2]
        self.S_bk_OCBS_phi = self.k ** 2 * s2 * o[:, 3]
        self.S_bk_OCBS = (-self.S_bk_OCBS_theta + self.S_bk_OCBS_phi) / s2
        return self.S_bk_OCBS
