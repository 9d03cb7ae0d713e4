"""GRE (Gaussian Random Ellipsoid) shape model and voxeliser — oracle.

Restates ``shape_model/gaussian_ellipsoid.py`` of the reference (class
``gaussian_ellipsoid_shape_model``, lines 44-237 and 306-310) and
``run_dda._build_gre_geometry`` (``run_dda.py:21-35``).  The floating-point
operation order of every step that feeds the voxel mask is kept as in the
reference, so that the mask is bit-identical for the same surface points.

Pinned by: the ``test_dda.ipynb`` recorded ``lattice_n = [17,17,17]`` and Mie
values (which fix N_occ through r_ve), the 4 ``r_ve`` values stored in
``dda_results/pcas_ocbs_simulated_data.hdf5`` (beta = 0 oblate spheroids,
bc = 3), closed-form ellipsoid checks, and brute-force nearest distances
against the KD-tree on small inputs.  For beta > 0 the surface depends on
LAPACK ``eigh`` and FITPACK rounding of the installed NumPy/SciPy; masks for
beta > 0 are pinned only against this module (parity unpinned against the
reference's pinned library versions).
"""
import numpy as np
import scipy.interpolate
from sklearn.neighbors import KDTree


class GREShape:
    """Shape parameters and derived lattice spacing.

    ``gaussian_ellipsoid.py:52-70``: c from the volume-equivalent radius,
    spacing ``lattice_lf`` and the near-surface threshold ``distance_factor``.
    """

    def __init__(self, r_v_base, bc_ratio, ab_ratio, beta):
        self.r_v_base = r_v_base
        self.bc_ratio = bc_ratio
        self.ab_ratio = ab_ratio
        self.beta = beta
        self.name = None
        c = np.cbrt(r_v_base ** 3 / (ab_ratio * bc_ratio ** 2))
        self.lattice_lf = (1 / 75) * np.sqrt(c / 0.05) * (ab_ratio * bc_ratio) ** (1 / 5)
        self.distance_factor = 3 ** 0.5 / 2 * (bc_ratio * ab_ratio * (c / 0.1)) ** (1 / 6)

    def semi_axes(self):
        """(a, b, c, h0, lc) as in ``gaussian_ellipsoid.py:90-94``."""
        c = np.cbrt(self.r_v_base ** 3 / (self.ab_ratio * self.bc_ratio ** 2))
        b = c * self.bc_ratio
        a = b * self.ab_ratio
        return a, b, c, c ** 2 / a, c * 0.3

    def coarse_grid(self):
        """Coarse (theta, phi) grid and its ellipsoid points, ``gaussian_ellipsoid.py:96-104``."""
        a, b, c, _, _ = self.semi_axes()
        th = np.linspace(0, np.pi, 25)
        ph = np.linspace(0, 2 * np.pi, 100)
        thm, phm = np.meshgrid(th, ph, indexing='ij')
        px = (a * np.sin(thm) * np.cos(phm)).ravel()
        py = (b * np.sin(thm) * np.sin(phm)).ravel()
        pz = (c * np.cos(thm)).ravel()
        return th, ph, px, py, pz

    def covariance(self):
        """beta^2 exp(-d^2 / (2 lc^2)) on the coarse grid with component-wise
        squared differences summed x, then y, then z (``:108-113``)."""
        _, _, _, _, lc = self.semi_axes()
        _, _, px, py, pz = self.coarse_grid()
        d2 = ((px[:, None] - px[None, :]) ** 2
              + (py[:, None] - py[None, :]) ** 2
              + (pz[:, None] - pz[None, :]) ** 2)
        return self.beta ** 2 * np.exp(-0.5 * d2 / lc ** 2)

    def coarse_field(self, rng):
        """One ``multivariate_normal(method='eigh')`` draw of the log-radius
        field on the 25 x 100 grid (``:115-120``)."""
        cov = self.covariance()
        return rng.multivariate_normal(np.zeros(cov.shape[0]), cov, size=1,
                                       method='eigh').squeeze().reshape(25, 100)

    def surface_from_field(self, s):
        """Deformed surface points for a coarse field s (25 x 100): bicubic
        spline upsampling by int(4 cbrt(bc*ab)) (``:122-130``) and normal
        displacement h0((e^s - beta^2/2) - 1) (``:132-147``).
        Returns (P, 3) points, P = (25 f) * (100 f)."""
        a, b, c, h0, _ = self.semi_axes()
        th, ph, _, _, _ = self.coarse_grid()
        spline = scipy.interpolate.RectBivariateSpline(th, ph, s)
        fac = int(4 * np.cbrt(self.bc_ratio * self.ab_ratio))
        th2 = np.linspace(0, np.pi, fac * 25)
        ph2 = np.linspace(0, 2 * np.pi, fac * 100)
        thm2, phm2 = np.meshgrid(th2, ph2, indexing='ij')
        s2 = spline(th2, ph2)
        x0 = a * np.sin(thm2) * np.cos(phm2)
        y0 = b * np.sin(thm2) * np.sin(phm2)
        z0 = c * np.cos(thm2)
        nrm = np.sqrt((x0 / a ** 2) ** 2 + (y0 / b ** 2) ** 2 + (z0 / c ** 2) ** 2)
        nvec = np.array([x0 / a ** 2, y0 / b ** 2, z0 / c ** 2]) / nrm
        deform = h0 * ((np.exp(s2) - 0.5 * self.beta ** 2) - 1)
        xg = x0 + deform * nvec[0]
        yg = y0 + deform * nvec[1]
        zg = z0 + deform * nvec[2]
        return np.column_stack([xg.ravel(), yg.ravel(), zg.ravel()])

    def surface_points(self, rng):
        """Deformed surface points, ``gaussian_ellipsoid.py:73-147``: the
        coarse field draw, then the spline and the normal displacement."""
        self.name = ("GRE_shape__r_v_base={:.2f}um__ab_ratio={:.1f}"
                     "__bc_ratio={:.1f}__beta={:.2f}".format(
                         self.r_v_base, self.ab_ratio, self.bc_ratio, self.beta))
        return self.surface_from_field(self.coarse_field(rng))

    def cuboid_lattice(self, pts):
        """Enclosing cuboid lattice, ``gaussian_ellipsoid.py:150-179``.

        Per axis ``np.arange(-lim, lim, lf)`` with lim = max|coord| + lf;
        points flattened from an 'ij' meshgrid (address = ix*Ny*Nz+iy*Nz+iz).
        Returns (lattice_n int32[3], grid_points (n, 3)).
        """
        lf = self.lattice_lf
        arrs = [np.arange(-(np.max(np.abs(pts[:, d])) + lf),
                          np.max(np.abs(pts[:, d])) + lf, lf) for d in range(3)]
        n = np.array([len(a) for a in arrs], dtype=np.int32)
        g = np.meshgrid(*arrs, indexing='ij')
        return n, np.column_stack([g[0].ravel(), g[1].ravel(), g[2].ravel()])


def nearest_distance(grid_points, surf_points):
    """Nearest surface-point distance, ``gaussian_ellipsoid.py:182-192`` (KD-tree)."""
    dist, _ = KDTree(surf_points).query(grid_points, k=1)
    return dist[:, 0]


def nearest_distance_brute(grid_points, surf_points, chunk=2048):
    """Brute-force nearest distance with the KD-tree's per-pair formula.

    sqrt(((dx*dx) + dy*dy) + dz*dz) minimised over all surface points; the
    minimum is order-free, so this equals the KD-tree result bit for bit.
    """
    out = np.empty(len(grid_points))
    for s in range(0, len(grid_points), chunk):
        g = grid_points[s:s + chunk]
        dx = g[:, None, 0] - surf_points[None, :, 0]
        dy = g[:, None, 1] - surf_points[None, :, 1]
        dz = g[:, None, 2] - surf_points[None, :, 2]
        r2 = dx * dx
        r2 += dy * dy
        r2 += dz * dz
        out[s:s + chunk] = np.sqrt(r2.min(axis=1))
    return out


def inside_mask(lattice_lf, distance_factor, lattice_n, dist):
    """Three-axis ray-cast inside test, ``gaussian_ellipsoid.py:195-237``.

    A voxel is near the surface if dist < distance_factor*lf (``:219``); it is
    inside along an axis if at least one near voxel lies strictly before it and
    one strictly after it (cumulative counts, ``:223-235``); AND over 3 axes.
    """
    nx, ny, nz = (int(v) for v in lattice_n)
    near = (dist.reshape(nx, ny, nz) < distance_factor * lattice_lf).astype(np.int32)
    ins = np.ones((nx, ny, nz), dtype=bool)
    for ax in (2, 1, 0):
        fwd = np.cumsum(near, axis=ax)
        rev = np.flip(np.cumsum(np.flip(near, axis=ax), axis=ax), axis=ax)
        ins &= ((fwd - near) >= 1) & ((rev - near) >= 1)
    return ins.ravel()


def volume_and_ve_radius(lattice_lf, mask):
    """``gaussian_ellipsoid.py:306-310``."""
    vol = lattice_lf ** 3 * np.sum(mask)
    return vol, np.cbrt(3 * vol / (4 * np.pi))


def build_gre_geometry(r_v_base, bc_ratio, ab_ratio, beta, seed=12345, brute=False):
    """``run_dda._build_gre_geometry`` (``run_dda.py:21-35``).

    Returns (name, lattice_n, lattice_lf, grid_points, is_in, r_ve, surf_pts).
    """
    rng = np.random.default_rng(seed)
    g = GREShape(r_v_base, bc_ratio, ab_ratio, beta)
    pts = g.surface_points(rng)
    n, grid = g.cuboid_lattice(pts)
    dist = (nearest_distance_brute if brute else nearest_distance)(grid, pts)
    is_in = inside_mask(g.lattice_lf, g.distance_factor, n, dist)
    r_ve = np.cbrt(3.0 * g.lattice_lf ** 3 * int(is_in.sum()) / (4 * np.pi))
    return g.name, n, g.lattice_lf, grid, is_in, r_ve, pts
