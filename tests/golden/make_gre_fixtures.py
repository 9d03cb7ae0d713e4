"""Write tests/golden/gre_fields.npz from the reference's GRE shape model.

Runs only in a container where the reference tree is mounted (read-only)
at /root/reference (override with BDDA_REFERENCE).  It imports the
reference's ``shape_model/gaussian_ellipsoid.py`` (class
``gaussian_ellipsoid_shape_model``, lines 73-237) and runs it on the GRE
particles of the benchmark workloads (the 24^3 and 96^3 isotropic GREs, the
64^3 birefringent GRE and the 32 batched instances), recording for each:

* the covariance matrix handed to ``rng.multivariate_normal`` (its SHA-256:
  pins lc and the squared-distance order, ``:106-113``),
* the coarse log-radius field ``s`` (25 x 100) given to
  ``RectBivariateSpline`` (``:120-122``),
* the surface points (SHA-256 of the (P, 3) float64 array, plus every 37th
  point), the lattice dims and half-widths (``:164-174``), and
* the voxel mask of ``extract_lattice_address_in_GRE_volume`` (``:195-237``)
  on the KD-tree distances (``:182-192``), packed 8 cells a byte.

Only data is written; nothing of the reference's source travels.  The
covariance eigendecomposition (``method='eigh'``) is LAPACK's, so ``s``
records this host's draw (with the SHA-256 of this host's ``eigh`` of the
covariance, so a test can tell whether its own LAPACK gives the same draw);
every consumer rebuilds the surface from ``s``.

    python tests/golden/make_gre_fixtures.py
"""
import hashlib
import importlib.util
import os
import sys

import numpy as np
import scipy.interpolate

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("BDDA_REFERENCE", "/root/reference")
OUT = os.path.join(HERE, "gre_fields.npz")


def load_reference_model():
    os.environ.setdefault("MPLBACKEND", "Agg")
    path = os.path.join(REF, "shape_model", "gaussian_ellipsoid.py")
    spec = importlib.util.spec_from_file_location("ref_gaussian_ellipsoid", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.gaussian_ellipsoid_shape_model


class RecordingRng:
    """Wraps a numpy Generator and records the covariance of its draw."""

    def __init__(self, rng):
        self._rng = rng
        self.cov = None

    def multivariate_normal(self, mean, cov, **kw):
        self.cov = np.array(cov, copy=True)
        return self._rng.multivariate_normal(mean, cov, **kw)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def workloads():
    sys.path.insert(0, os.path.join(ROOT, "block-dda_py_amd"))
    from bdda_amd import workloads as w
    out = [w.SMALL_24, w.FULL_64, w.LARGE_96]
    out += w.batch_instances(w.BATCH_32)
    return out


def main():
    Model = load_reference_model()
    fields = {}
    real_spline = scipy.interpolate.RectBivariateSpline
    seen = {}

    class RecordingSpline(real_spline):
        def __init__(self, x, y, z, *a, **kw):
            seen["s"] = np.array(z, copy=True)
            super().__init__(x, y, z, *a, **kw)

    scipy.interpolate.RectBivariateSpline = RecordingSpline
    try:
        for wl in workloads():
            g = Model(wl["r_v_base"], wl["bc_ratio"], wl["ab_ratio"], wl["beta"])
            rng = RecordingRng(np.random.default_rng(wl["shape_seed"]))
            pts, _ = g.compute_r_points_on_GRE(rng)
            domain, n, grid = g.create_cuboid_lattice_that_encloses_GRE_shape(pts)
            dist = g.find_nearest_distance_from_the_GRE_surf(grid, pts)
            is_in = Model.extract_lattice_address_in_GRE_volume(g.lattice_lf, g.distance_factor, n, dist)
            key = wl["name"]
            fields[key + "__params"] = np.array([wl["r_v_base"], wl["bc_ratio"], wl["ab_ratio"], wl["beta"],
                                                wl["shape_seed"]], dtype=np.float64)
            fields[key + "__s"] = seen.pop("s")
            fields[key + "__cov_sha256"] = np.array(sha(rng.cov))
            fields[key + "__pts_sha256"] = np.array(sha(pts))
            fields[key + "__pts_every37"] = pts[::37].copy()
            fields[key + "__lattice_n"] = np.asarray(n, dtype=np.int32)
            fields[key + "__lims"] = np.array([d[1] for d in domain], dtype=np.float64)
            fields[key + "__mask_bits"] = np.packbits(is_in.astype(np.uint8))
            fields[key + "__n_occ"] = np.array(int(is_in.sum()), dtype=np.int64)
            fields[key + "__dist_sha256"] = np.array(sha(dist))
            # this host's eigendecomposition of the covariance: a host whose
            # LAPACK returns the same eigenvectors reproduces the draw itself
            w, v = np.linalg.eigh(rng.cov)
            fields[key + "__eigh_sha256"] = np.array(sha(w) + sha(v))
            print(f"{key}: lattice {tuple(int(v) for v in n)}, {int(is_in.sum())} occupied", flush=True)
    finally:
        scipy.interpolate.RectBivariateSpline = real_spline
    fields["names"] = np.array([k[:-len("__s")] for k in fields if k.endswith("__s")])
    np.savez_compressed(OUT, **fields)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
