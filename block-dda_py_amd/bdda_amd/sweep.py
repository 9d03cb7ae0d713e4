"""Parameter sweep over GRE shapes, (wavelength, medium) pairs and particle
refractive indices, writing the reference's results layout.

Mirrors the reference's sweep driver ``run_dda.py`` (lines 10-146) and the
file layout that ``dda_results/create_h5py.ipynb`` (cell 1) creates:

* geometry built once per shape with ``default_rng(RNG_SEED)``
  (``run_dda.py:21-35``), its r_ve stored per shape;
* per condition, a solve with orientations from ``default_rng(RNG_SEED + 1)``
  drawn alpha, beta, gamma per try; up to MAX_TRY tries with fresh
  orientations while the solver does not converge, then NaN fill
  (``:38-70``);
* a condition whose stored Mie S(0) has a nonzero imaginary part is already
  done and skipped (resume, ``:102-105``);
* the Mie values of the volume-equivalent sphere with the axis-averaged m_p
  (``:123-128``);
* datasets (N_pairs, N_m_p, N_rv, N_bc, N_ab, N_bt[, N_ori[, 3]]) under
  ``target/simulated_data`` with their ``definition`` attributes.

What changes on the GPU: the m_p conditions of one (shape, wavelength,
medium) share one Green spectrum (A_off does not depend on m_p), and their
first tries -- which use the same orientations -- run as one batch of
concurrent solves (``bdda_solve_batch``; per-instance results equal to
one-at-a-time solves).  The file is an HDF5 file when h5py is importable
and the path ends in .h5/.hdf5; otherwise an .npz file with the same dataset
paths (``target/...``) and attributes (stored as JSON), rewritten after each
condition so that an interrupted sweep resumes where it stopped.
"""
import datetime
import itertools
import json
import os

import numpy as np

RNG_SEED = 12345
MAX_TRY = 4

_SD = "target/simulated_data/"


def _log_default(msg):
    print(f"[{datetime.datetime.now():%H:%M:%S}] {msg}", flush=True)


# (name, dims after the condition axes, dtype, definition) -- create_h5py.ipynb cell 1
_DATASETS = [
    ("r_ve", "geo", np.float64, "volume-equivalent radius [um] from discretised particle volume"),
    ("Euler_angles", "ang", np.float64,
     "ZYZ Euler angles (alpha,beta,gamma) rotating particle frame from lab frame [rad]"),
    ("C_abs", "ori", np.float64, "absorption cross section per orientation [um^2]"),
    ("C_ext", "ori", np.float64, "extinction cross section per orientation [um^2]"),
    ("S_fw_PCAS_theta", "ori", np.complex128, "forward-scattering amplitude S(0)_theta per orientation [um]"),
    ("S_fw_PCAS_phi", "ori", np.complex128, "forward-scattering amplitude S(0)_phi per orientation [um]"),
    ("S_bk_OCBS", "ori", np.complex128, "backward-scattering amplitude S(180) per orientation [um]"),
    ("C_abs_mie", "cond", np.float64, "Mie C_abs of volume-equivalent sphere [um^2]"),
    ("C_ext_mie", "cond", np.float64, "Mie C_ext of volume-equivalent sphere [um^2]"),
    ("S_fw_PCAS_mie", "cond", np.complex128, "Mie S(0) of volume-equivalent sphere [um]"),
    ("S_bk_OCBS_mie", "cond", np.complex128, "Mie S(180) of volume-equivalent sphere [um]"),
]

_TARGET_ATTRS = {
    "light_source": "Thorlabs SLD830S-A10 @ I=140mA, T=25C",
    "polarization_state": "left-handed circular: E0_theta=1/sqrt(2), E0_phi=1j/sqrt(2)",
}
_SD_ATTRS = {
    "scattering_code": "block-DDA_Py (discrete dipole approximation)",
    "orientation": "random ZYZ Euler angles: alpha in [0,2pi), beta in [0,pi), gamma in [0,2pi)",
    "units": "r_ve:[um], euler_angles:[rad], C:[um^2], S:[um]",
    "S_definition": ("S(0)_theta = S11(0)+1j*S12(0), S(0)_phi = S22(0)-1j*S21(0), "
                     "S(180) = (S11+S22+1j*S12-1j*S21)(180)/sqrt(2), per Mishchenko 2000"),
}


class NpzStore:
    """Datasets keyed by their HDF5 path ('target/...'), attributes per path,
    kept in memory and written to an .npz file by flush()."""

    def __init__(self, path, arrays, attrs):
        self.path = path
        self.arrays = arrays
        self.attrs = attrs

    def __getitem__(self, key):
        return self.arrays[key]

    def __contains__(self, key):
        return key in self.arrays

    def flush(self):
        tmp = self.path + ".tmp.npz"
        np.savez(tmp, __attrs__=np.array(json.dumps(self.attrs)), **self.arrays)
        os.replace(tmp, self.path)

    def close(self):
        self.flush()

    @classmethod
    def open(cls, path):
        with np.load(path, allow_pickle=False) as d:
            arrays = {k: d[k].copy() for k in d.files if k != "__attrs__"}
            attrs = json.loads(str(d["__attrs__"])) if "__attrs__" in d.files else {}
        return cls(path, arrays, attrs)


class H5Store:
    """The same interface over an h5py file (the reference's format)."""

    def __init__(self, f):
        self.f = f
        self.path = f.filename

    def __getitem__(self, key):
        return self.f[key]

    def __contains__(self, key):
        return key in self.f

    @property
    def attrs(self):
        return {k: dict(self.f[k].attrs) for k in ("target", "target/simulated_data")}

    def flush(self):
        self.f.flush()

    def close(self):
        self.f.close()


def _use_h5(path):
    if not path.endswith((".h5", ".hdf5")):
        return False
    try:
        import h5py  # noqa: F401
        return True
    except ImportError:
        return False


def create_store(path, wl_m_m_pairs, m_p_xyz_list, r_v_base_list, bc_ratio_list, ab_ratio_list, gre_beta_list,
                 num_orientations):
    """A new results file with the parameter lists and zeroed datasets
    (create_h5py.ipynb cell 1).  Returns the store."""
    lists = {
        "target/wl_m_m_pairs": np.asarray(wl_m_m_pairs, dtype=np.float64).reshape(-1, 2),
        "target/m_p_xyz_list": np.asarray(m_p_xyz_list, dtype=np.complex128).reshape(-1, 3),
        "target/r_v_base_list": np.asarray(r_v_base_list, dtype=np.float64).ravel(),
        "target/bc_ratio_list": np.asarray(bc_ratio_list, dtype=np.float64).ravel(),
        "target/ab_ratio_list": np.asarray(ab_ratio_list, dtype=np.float64).ravel(),
        "target/gre_beta_list": np.asarray(gre_beta_list, dtype=np.float64).ravel(),
    }
    geo = tuple(len(lists["target/" + k]) for k in ("r_v_base_list", "bc_ratio_list", "ab_ratio_list",
                                                    "gre_beta_list"))
    cond = (len(lists["target/wl_m_m_pairs"]), len(lists["target/m_p_xyz_list"])) + geo
    shapes = {"geo": geo, "cond": cond, "ori": cond + (int(num_orientations),),
              "ang": cond + (int(num_orientations), 3)}
    tattrs = dict(_TARGET_ATTRS, num_orientations=int(num_orientations))
    if _use_h5(path):
        import h5py
        f = h5py.File(path, "w")
        g = f.create_group("target")
        g.attrs.update(tattrs)
        for k, v in lists.items():
            f.create_dataset(k, data=v)
        sd = g.create_group("simulated_data")
        sd.attrs.update(_SD_ATTRS)
        for name, kind, dt, definition in _DATASETS:
            d = sd.create_dataset(name, shapes[kind], dtype=dt)
            d.attrs["definition"] = definition
        return H5Store(f)
    arrays = dict(lists)
    attrs = {"target": tattrs, "target/simulated_data": dict(_SD_ATTRS)}
    for name, kind, dt, definition in _DATASETS:
        arrays[_SD + name] = np.zeros(shapes[kind], dtype=dt)
        attrs[_SD + name] = {"definition": definition}
    st = NpzStore(path, arrays, attrs)
    st.flush()
    return st


def open_store(path):
    """An existing results file (HDF5 opened r+, or .npz)."""
    if _use_h5(path):
        import h5py
        return H5Store(h5py.File(path, "r+"))
    return NpzStore.open(path)


def build_gre_geometry(r_v_base, bc_ratio, ab_ratio, gre_beta, seed=RNG_SEED):
    """(name, lattice_n, lattice_lf, grid, is_in, r_ve) of one shape, from
    default_rng(seed) (run_dda.py:21-35); nearest distances and the inside
    test on the GPU."""
    from shape_model.gaussian_ellipsoid import gaussian_ellipsoid_shape_model as G
    rng = np.random.default_rng(seed)
    g = G(r_v_base, bc_ratio, ab_ratio, gre_beta)
    pts, _ = g.compute_r_points_on_GRE(rng)
    _, lattice_n, grid = g.create_cuboid_lattice_that_encloses_GRE_shape(pts)
    dist = g.find_nearest_distance_from_the_GRE_surf(grid, pts)
    is_in = g.extract_lattice_address_in_GRE_volume(g.lattice_lf, g.distance_factor, lattice_n, dist)
    r_ve = np.cbrt(3.0 * g.lattice_lf ** 3 * int(is_in.sum()) / (4 * np.pi))
    return g.name, lattice_n, g.lattice_lf, grid, is_in, r_ve


def orientation_tries(num_orientations, max_try=MAX_TRY, seed=RNG_SEED + 1):
    """The Euler angles of tries 1..max_try: alpha, beta, gamma uniform,
    drawn per try from one default_rng(seed) (run_dda.py:46-54)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(max_try):
        out.append(np.column_stack([rng.uniform(0, 2 * np.pi, num_orientations),
                                    rng.uniform(0, np.pi, num_orientations),
                                    rng.uniform(0, 2 * np.pi, num_orientations)]))
    return out


def _outputs(dd):
    return (dd.compute_C_abs(), dd.compute_C_ext(), *dd.compute_PCAS_observable_S_fw(),
            dd.compute_OCBS_observable_S_bk())


def _nan_outputs(n):
    nan_r = np.full(n, np.nan)
    nan_c = np.full(n, np.nan + 0j)
    return nan_r, nan_r.copy(), nan_c, nan_c.copy(), nan_c.copy()


def _dipoles(target, wl_0, m_m, euler, options):
    from bl_dda.scatterer import DiscreteDipoles, IncidentField
    dd = DiscreteDipoles(target, IncidentField(wl_0, m_m, euler))
    for k, v in options.items():
        setattr(dd, k, v)
    dd.set_interaction_matrix()
    return dd


def run_dda(target, wl_0, m_m, num_orientations, max_try=MAX_TRY, seed=RNG_SEED + 1, log=_log_default,
            options=None, first=None):
    """One condition with retries (run_dda.py:38-70): returns (euler, C_abs,
    C_ext, S_fw_theta, S_fw_phi, S_bk, converged, tries).  `first`, if given,
    is an already solved DiscreteDipoles of try 1."""
    options = options or {}
    tries = orientation_tries(num_orientations, max_try, seed)
    euler = None
    for t, euler in enumerate(tries, start=1):
        if t == 1 and first is not None:
            dd = first
        else:
            dd = _dipoles(target, wl_0, m_m, euler, options)
            dd.solve_matrix_equation()
        log(f"    try {t}/{max_try}: {'converged' if dd.converge else 'not converged'}")
        if dd.converge:
            return (euler, *_outputs(dd), True, t)
    return (euler, *_nan_outputs(num_orientations), False, max_try)


def run_sweep(store, rng_seed=RNG_SEED, max_try=MAX_TRY, log=_log_default, batch_m_p=True, nstreams=8,
              options=None):
    """Fill every condition of `store` not yet computed (run_dda.py:74-146).
    `options` sets DiscreteDipoles attributes (tol, itermax, solver,
    precision, ...); left empty, the reference's defaults apply (block-
    BiCGSTAB, tol 1e-2, itermax 25).  Returns a summary dict."""
    from analytical_scattering_theories.homogeneous_sphere import mie_compute_q_and_s
    from bl_dda.scatterer import DiscreteDipoles, Target
    from .batch import solve_batch
    options = dict(options or {})
    options.setdefault("verbose", False)
    t = "target/"
    num_orientations = int(store.attrs["target"]["num_orientations"])
    pairs = np.asarray(store[t + "wl_m_m_pairs"][:])
    m_p_list = np.asarray(store[t + "m_p_xyz_list"][:])
    lists = [np.asarray(store[t + k][:]) for k in ("r_v_base_list", "bc_ratio_list", "ab_ratio_list",
                                                   "gre_beta_list")]
    sd = {name: store[_SD + name] for name, _, _, _ in _DATASETS}
    summary = {"computed": 0, "skipped": 0, "not_converged": 0, "tries": []}
    first_euler = orientation_tries(num_orientations, 1, rng_seed + 1)[0]
    for (i_rv, r_v), (i_bc, bc), (i_ab, ab), (i_bt, bt) in itertools.product(*(enumerate(v) for v in lists)):
        shape4 = (i_rv, i_bc, i_ab, i_bt)
        name, lattice_n, lf, grid, is_in, r_ve = build_gre_geometry(r_v, bc, ab, bt, rng_seed)
        sd["r_ve"][shape4] = r_ve
        for i_pair, (wl_0, m_m) in enumerate(pairs):
            todo = []
            for i_mp, m_p_xyz in enumerate(m_p_list):
                idx6 = (i_pair, i_mp) + shape4
                if np.asarray(sd["S_fw_PCAS_mie"][idx6]).imag != 0.0:
                    log(f"Skip: pair={i_pair} m_p={i_mp} shape={shape4} (already computed)")
                    summary["skipped"] += 1
                    continue
                todo.append((i_mp, m_p_xyz, idx6))
            if not todo:
                continue
            targets = [Target(name, lattice_n, lf, grid, is_in, m_p_xyz) for _, m_p_xyz, _ in todo]
            firsts = [None] * len(todo)
            if batch_m_p and len(todo) > 1:
                # try 1 of every m_p uses the same orientations: one batch
                dds = [_dipoles(tg, wl_0, m_m, first_euler, options) for tg in targets]
                solve_batch(dds, nstreams=nstreams)
                firsts = dds
            for (i_mp, m_p_xyz, idx6), tg, first in zip(todo, targets, firsts):
                log(f"wl_0={wl_0:.4f} um  m_m={m_m:.4f}  m_p_xyz={m_p_xyz}  |  r_v_base={r_v:.3f}  bc={bc:.1f}  "
                    f"ab={ab:.1f}  beta={bt:.2f}  r_ve={r_ve:.4f} um  N_ori={num_orientations}")
                euler, c_abs, c_ext, s_th, s_ph, s_bk, ok, tries = run_dda(
                    tg, wl_0, m_m, num_orientations, max_try, rng_seed + 1, log, options, first)
                m_p_avg = complex(np.mean(m_p_xyz))
                _, q_abs, q_ext, s_fw_mie, s_bk_mie = mie_compute_q_and_s(wl_0, m_m, r_ve, m_p_avg, nang=3)
                sl = slice(None)
                sd["Euler_angles"][idx6 + (sl, sl)] = euler
                sd["C_abs"][idx6 + (sl,)] = c_abs
                sd["C_ext"][idx6 + (sl,)] = c_ext
                sd["S_fw_PCAS_theta"][idx6 + (sl,)] = s_th
                sd["S_fw_PCAS_phi"][idx6 + (sl,)] = s_ph
                sd["S_bk_OCBS"][idx6 + (sl,)] = s_bk
                sd["C_abs_mie"][idx6] = q_abs * np.pi * r_ve ** 2
                sd["C_ext_mie"][idx6] = q_ext * np.pi * r_ve ** 2
                sd["S_fw_PCAS_mie"][idx6] = s_fw_mie
                sd["S_bk_OCBS_mie"][idx6] = s_bk_mie
                store.flush()
                summary["computed"] += 1
                summary["not_converged"] += int(not ok)
                summary["tries"].append(tries)
                log(f"  C_ext(mean)={np.nanmean(c_ext):.4e}  S_fw_theta(mean)={np.nanmean(s_th):.4g}  "
                    f"S_fw_phi(mean)={np.nanmean(s_ph):.4g}  S_bk(mean)={np.nanmean(s_bk):.4g}")
    return summary
