"""Synthetic benchmark workloads (GRE particles, orientations, materials).

Each workload is a plain dict of shape, material and solver settings; the
geometry is built by the drop-in GRE model (host surface sampling, GPU
voxeliser) and the orientations follow the reference's draw order
(``run_dda.py:46-54``: alpha, beta, gamma uniform from one generator).
"""
import os

import numpy as np

# 64^3 cuboid: GRE bc = 0.8, ab = 0.7, beta = 0.2 (seed 0).  r_v gives a
# 51 x 48 x 63 lattice (21,113 dipoles, from the recorded field) padded with
# vacuum to 64^3, i.e. a 128^3 doubled FFT grid; birefringent m_p; L = 64;
# BiCGSTAB to 1e-5.
FULL_64 = dict(
    name="gre_bc0.8_ab0.7_beta0.2_seed0_64cube_L64",
    r_v_base=1.0107600709889086, bc_ratio=0.8, ab_ratio=0.7, beta=0.2, shape_seed=0,
    m_p_xyz=(1.55 + 0.001j, 1.60 + 0.001j, 1.50 + 0.001j), m_m=1.33, dpl=15.0,
    L=64, orient_seed=1, pad_to=(64, 64, 64), solver="bicgstab", tol=1e-5, itermax=500,
)

# 24^3 isotropic sphere-like GRE (beta = 0.1, seed 0), m_p = 1.5 + 0.01i,
# m_m = 1, size parameter ~3, L = 8, COCG-rq to 1e-5.
SMALL_24 = dict(
    name="gre_bc1_ab1_beta0.1_seed0_24cube_L8",
    r_v_base=0.35858192704617975, bc_ratio=1.0, ab_ratio=1.0, beta=0.1, shape_seed=0,
    m_p_xyz=(1.5 + 0.01j,) * 3, m_m=1.0, size_parameter=3.0,
    L=8, orient_seed=1, pad_to=None, solver="cocg_rq", tol=1e-5, itermax=500,
)


# Config-1 physics (bc = ab = 1, beta = 0.1, seed 0, m_p = 1.5 + 0.01i, m_m = 1)
# scaled to a 96^3 cuboid at the same resolution: r_v = 6.2 um gives a
# 91 x 92 x 87 lattice (307,817 dipoles, lf = 0.148 um) padded to 96^3
# (192^3 FFT grid),
# size parameter 12 (about 16 dipoles per wavelength, as config 1's 14);
# L = 32; mixed-precision block-COCG-rq to 1e-5.
LARGE_96 = dict(
    name="gre_bc1_ab1_beta0.1_seed0_96cube_L32",
    r_v_base=6.2, bc_ratio=1.0, ab_ratio=1.0, beta=0.1, shape_seed=0,
    m_p_xyz=(1.5 + 0.01j,) * 3, m_m=1.0, size_parameter=12.0,
    L=32, orient_seed=1, pad_to=(96, 96, 96), solver="cocg_rq", tol=1e-5, itermax=500,
)

# Batched uneven instances: 32 GRE particles (shape seeds 0-31) whose longest
# lattice axis is drawn uniformly from 20..48 cells, bc and ab uniform in
# [0.5, 1], beta uniform in [0, 0.3], isotropic m_p with real part uniform in
# [1.4, 1.7] and imaginary part log-uniform in [1e-4, 1e-1], m_m = 1,
# 15 dipoles per wavelength, L = 16 orientations each (orientation seed =
# shape seed + 1); mixed-precision block-COCG-rq to 1e-5.
BATCH_32 = dict(
    name="batch32_gre_grid20to48_L16", count=32, grid=(20, 48), bc=(0.5, 1.0), ab=(0.5, 1.0), beta=(0.0, 0.3),
    re_m=(1.4, 1.7), im_m=(1e-4, 1e-1), m_m=1.0, dpl=15.0, L=16, solver="cocg_rq", precision="mixed", tol=1e-5,
    itermax=500,
)


def rv_for_grid(cells, bc, ab):
    """r_v whose beta = 0 ellipsoid spans about `cells` lattice cells along
    its longest axis: the axis is 2c, the spacing (1/75) sqrt(c/0.05)
    (ab bc)^(1/5), and c = cbrt(r_v^3 / (ab bc^2))."""
    c = ((cells - 2) * (ab * bc) ** 0.2 / 150.0) ** 2 / 0.05
    return float(np.cbrt(c ** 3 * ab * bc ** 2))


def batch_instances(wl, count=None):
    """Per-instance workload dicts of a batched workload: instance i draws its
    size, shape and material from default_rng(i) and uses shape seed i."""
    out = []
    for i in range(wl["count"] if count is None else count):
        r = np.random.default_rng(i)
        cells = int(r.integers(wl["grid"][0], wl["grid"][1] + 1))
        bc, ab = r.uniform(*wl["bc"]), r.uniform(*wl["ab"])
        beta = r.uniform(*wl["beta"])
        m = r.uniform(*wl["re_m"]) + 1j * np.exp(r.uniform(np.log(wl["im_m"][0]), np.log(wl["im_m"][1])))
        out.append(dict(name=f"{wl['name']}_{i}", r_v_base=rv_for_grid(cells, bc, ab), bc_ratio=bc, ab_ratio=ab,
                        beta=beta, shape_seed=i, m_p_xyz=(m,) * 3, m_m=wl["m_m"], dpl=wl["dpl"], L=wl["L"],
                        orient_seed=i + 1, pad_to=None, solver=wl["solver"], tol=wl["tol"],
                        itermax=wl["itermax"]))
    return out


def euler_angles(L, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(0, 2 * np.pi, L), rng.uniform(0, np.pi, L), rng.uniform(0, 2 * np.pi, L)])


# Coarse GRE log-radius fields of the workloads above, drawn once on an
# x86-64 host with OpenBLAS (see record_fields).  The draw goes through
# LAPACK's eigendecomposition of the covariance, whose eigenvector signs
# depend on the BLAS kernel the CPU selects, so the same seed gives another
# particle on another host; the recorded fields make every host build the
# same particles.
FIELDS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "gre_fields.npz")
_fields = None


def recorded_field(wl):
    """The recorded 25 x 100 coarse field of workload `wl`, or None."""
    global _fields
    if _fields is None:
        _fields = {}
        if os.path.exists(FIELDS_PATH):
            with np.load(FIELDS_PATH) as d:
                _fields = {k: d[k] for k in d.files}
    return _fields.get(wl["name"])


def all_workloads():
    """Every GRE particle the workloads use (the batch expanded)."""
    return [SMALL_24, FULL_64, LARGE_96] + batch_instances(BATCH_32)


def record_fields(path=FIELDS_PATH):
    """Draw and store the coarse field of every workload particle
    (gaussian_ellipsoid.py:106-120: one multivariate_normal draw with the
    workload's shape seed)."""
    from shape_model.gaussian_ellipsoid import gaussian_ellipsoid_shape_model as G
    out = {}
    for wl in all_workloads():
        g = G(wl["r_v_base"], wl["bc_ratio"], wl["ab_ratio"], wl["beta"])
        out[wl["name"]] = g.draw_field(np.random.default_rng(wl["shape_seed"]))
    np.savez_compressed(path, **out)
    return path


def build_geometry(wl, recorded=True):
    """(name, lattice_n, lf, grid_points, is_in) via the GPU voxeliser.
    With `recorded`, a workload whose coarse field is recorded is built from
    it; otherwise (or if none is) the field is drawn from its shape seed."""
    from shape_model.gaussian_ellipsoid import gaussian_ellipsoid_shape_model as G
    g = G(wl["r_v_base"], wl["bc_ratio"], wl["ab_ratio"], wl["beta"])
    s = recorded_field(wl) if recorded else None
    if s is None:
        pts, _ = g.compute_r_points_on_GRE(np.random.default_rng(wl["shape_seed"]))
    else:
        g.set_name()
        pts, _ = g.surface_from_sample(s)
    _, n, grid = g.create_cuboid_lattice_that_encloses_GRE_shape(pts)
    dist = g.find_nearest_distance_from_the_GRE_surf(grid, pts)
    is_in = g.extract_lattice_address_in_GRE_volume(g.lattice_lf, g.distance_factor, n, dist)
    return g.name, n, g.lattice_lf, grid, is_in


def wavelength(wl, lf, n_occ):
    """Vacuum wavelength from the workload's dpl or size parameter."""
    m_p = np.asarray(wl["m_p_xyz"])
    if "dpl" in wl:
        return wl["dpl"] * lf * np.abs(np.max(m_p))
    r_ve = np.cbrt(3 * lf ** 3 * n_occ / (4 * np.pi))
    return 2 * np.pi * wl["m_m"] * r_ve / wl["size_parameter"]


def make_dipoles(wl, geometry=None, precision="fp64", solver=None, verbose=False, euler=None):
    """DiscreteDipoles for a workload, interaction matrix and B set on the
    device; `euler` (L, 3) replaces the workload's own orientations."""
    from bl_dda.scatterer import DiscreteDipoles, IncidentField, Target
    name, n, lf, grid, is_in = geometry if geometry is not None else build_geometry(wl)
    m_p = np.array(wl["m_p_xyz"], dtype=np.complex128)
    tg = Target(name, n, lf, grid, is_in, m_p)
    wl0 = wavelength(wl, lf, tg.num_element_occupy)
    if euler is None:
        euler = euler_angles(wl["L"], wl["orient_seed"])
    inc = IncidentField(wl0, wl["m_m"], euler)
    dd = DiscreteDipoles(tg, inc)
    dd.verbose = verbose
    dd.solver = solver or wl["solver"]
    dd.precision = precision
    dd.tol = wl["tol"]
    dd.itermax = wl["itermax"]
    dd.pad_to = wl["pad_to"]
    dd.set_interaction_matrix()
    return dd
