"""GRE (Gaussian Random Ellipsoid) shape model with a GPU voxeliser.

Drop-in for the reference ``shape_model/gaussian_ellipsoid.py`` class
``gaussian_ellipsoid_shape_model`` (v0.3.0, lines 44-237 and 306-310).
Surface sampling (one Gaussian draw on a 25 x 100 grid, spline upsampling)
and the enclosing lattice stay on the host, as in the reference.  The two
per-lattice-point stages run on the GPU:

* ``find_nearest_distance_from_the_GRE_surf``: brute-force nearest surface
  point per lattice point (HIP), the same per-pair float64 formula as the
  KD-tree, so distances are bit-identical;
* ``extract_lattice_address_in_GRE_volume``: the three-axis ray-cast inside
  test (HIP), integer-exact.

The matplotlib viewer of the reference (``:240-303``) is not provided.
"""
import numpy as np
import scipy.interpolate

from bdda_amd.plan import inside_mask, nearest_distance


class gaussian_ellipsoid_shape_model:  # noqa: N801  (reference class name)
    """GRE lattice generator (Muinonen & Pieniluoma 2011)."""

    N_THETA, N_PHI = 25, 100

    def __init__(self, r_v_base, bc_ratio, ab_ratio, beta):
        self.r_v_base = r_v_base
        self.bc_ratio = bc_ratio
        self.ab_ratio = ab_ratio
        self.beta = beta
        self.name = None
        c = np.cbrt(self.r_v_base ** 3 / (self.ab_ratio * self.bc_ratio ** 2))
        self.lattice_lf = (1 / 75) * np.sqrt(c / 0.05) * (self.ab_ratio * self.bc_ratio) ** (1 / 5)
        self.distance_factor = 3 ** 0.5 / 2 * (bc_ratio * ab_ratio * (c / 0.1)) ** (1 / 6)

    def _axes(self):
        c = np.cbrt(self.r_v_base ** 3 / (self.ab_ratio * self.bc_ratio ** 2))
        b = c * self.bc_ratio
        a = b * self.ab_ratio
        return a, b, c

    def compute_r_points_on_GRE(self, rng):
        """Surface points and meshes of the deformed ellipsoid (gaussian_ellipsoid.py:73-147)."""
        self.set_name()
        return self.surface_from_sample(self.draw_field(rng))

    def set_name(self):
        """The particle's name string (gaussian_ellipsoid.py:86-88)."""
        self.name = ("GRE_shape__r_v_base={:.2f}um__ab_ratio={:.1f}"
                     "__bc_ratio={:.1f}__beta={:.2f}".format(self.r_v_base, self.ab_ratio, self.bc_ratio,
                                                            self.beta))

    def draw_field(self, rng):
        """One draw of the coarse log-radius field, N_THETA x N_PHI
        (gaussian_ellipsoid.py:106-120): covariance beta^2 exp(-d^2/(2 lc^2)),
        multivariate_normal with method='eigh'."""
        cov = self.beta ** 2 * self.correlation_matrix()
        draw = rng.multivariate_normal(np.zeros(self.N_THETA * self.N_PHI), cov, size=1, method='eigh')
        return draw.squeeze().reshape(self.N_THETA, self.N_PHI)

    def correlation_matrix(self):
        """exp(-d^2 / (2 lc^2)) on the coarse grid; independent of r_v and beta
        (all distances scale with c, as does lc = 0.3 c)."""
        a, b, c = self._axes()
        lc = c * 0.3
        th = np.linspace(0, np.pi, self.N_THETA)
        ph = np.linspace(0, 2 * np.pi, self.N_PHI)
        T, F = np.meshgrid(th, ph, indexing='ij')
        coarse = [(a * np.sin(T) * np.cos(F)).ravel(), (b * np.sin(T) * np.sin(F)).ravel(), (c * np.cos(T)).ravel()]
        # squared distances summed component by component (x, then y, then z)
        d2 = (coarse[0][:, None] - coarse[0][None, :]) ** 2
        d2 = d2 + (coarse[1][:, None] - coarse[1][None, :]) ** 2
        d2 = d2 + (coarse[2][:, None] - coarse[2][None, :]) ** 2
        return np.exp(-0.5 * d2 / lc ** 2)

    def surface_from_sample(self, s):
        """Surface points and meshes for a given coarse log-radius field s
        (N_THETA x N_PHI): spline upsampling and normal displacement
        (gaussian_ellipsoid.py:122-147)."""
        a, b, c = self._axes()
        h0 = c ** 2 / a
        th = np.linspace(0, np.pi, self.N_THETA)
        ph = np.linspace(0, 2 * np.pi, self.N_PHI)
        spline = scipy.interpolate.RectBivariateSpline(th, ph, s)
        up = int(4 * np.cbrt(self.bc_ratio * self.ab_ratio))
        th2 = np.linspace(0, np.pi, up * self.N_THETA)
        ph2 = np.linspace(0, 2 * np.pi, up * self.N_PHI)
        T2, F2 = np.meshgrid(th2, ph2, indexing='ij')
        s_up = spline(th2, ph2)
        x0 = a * np.sin(T2) * np.cos(F2)
        y0 = b * np.sin(T2) * np.sin(F2)
        z0 = c * np.cos(T2)
        norm = np.sqrt((x0 / a ** 2) ** 2 + (y0 / b ** 2) ** 2 + (z0 / c ** 2) ** 2)
        normal = np.array([x0 / a ** 2, y0 / b ** 2, z0 / c ** 2]) / norm
        shift = h0 * ((np.exp(s_up) - 0.5 * self.beta ** 2) - 1)
        xg, yg, zg = x0 + shift * normal[0], y0 + shift * normal[1], z0 + shift * normal[2]
        return np.column_stack([xg.ravel(), yg.ravel(), zg.ravel()]), (xg, yg, zg)

    def create_cuboid_lattice_that_encloses_GRE_shape(self, r_points_on_GRE_surf):
        """Enclosing cuboid lattice (gaussian_ellipsoid.py:150-179)."""
        lf = self.lattice_lf
        lims = [np.max(np.abs(r_points_on_GRE_surf[:, d])) + lf for d in range(3)]
        domain = tuple([-v, v] for v in lims)
        axes = [np.arange(-v, v, lf) for v in lims]
        lattice_n = np.array([len(v) for v in axes], dtype=np.int32)
        g = np.meshgrid(*axes, indexing='ij')
        return domain, lattice_n, np.column_stack([g[0].ravel(), g[1].ravel(), g[2].ravel()])

    def find_nearest_distance_from_the_GRE_surf(self, lattice_grid_points, r_points_on_GRE_surf):
        """Nearest surface-point distance per lattice point, on the GPU (gaussian_ellipsoid.py:182-192)."""
        return nearest_distance(lattice_grid_points, r_points_on_GRE_surf)

    @staticmethod
    def extract_lattice_address_in_GRE_volume(lattice_lf, distance_factor, lattice_n, dist_from_GRE):
        """Three-axis ray-cast inside test, on the GPU (gaussian_ellipsoid.py:195-237)."""
        return inside_mask(dist_from_GRE, lattice_n, distance_factor * lattice_lf)

    def compute_GRE_volume_and_ve_radius(self, lattice_grid_points_is_in_GREvol):
        """gaussian_ellipsoid.py:306-310."""
        num_in = np.sum(lattice_grid_points_is_in_GREvol)
        volume = self.lattice_lf ** 3 * num_in
        return volume, np.cbrt(3 * volume / (4 * np.pi))
