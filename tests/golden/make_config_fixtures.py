"""Write tests/golden/config_full64_L8.npz: the float64 oracle's solve of the
64^3 benchmark particle for its first 8 orientations.

The particle is the benchmark's GRE (bc 0.8, ab 0.7, beta 0.2, seed 0) built
from the coarse field the reference recorded (``gre_fields.npz``), on its
own lattice (the GPU solve pads it to a 64^3 cuboid, which leaves the exact
convolution unchanged).  Birefringent m_p = (1.55, 1.60, 1.50) + 0.001i,
m_m = 1.33, 15 dipoles per wavelength, block-BiCGSTAB to 1e-5
(``bl_krylov.py:121-187``), observables as ``scatterer.py:187-228``.  The
oracle takes about a minute here, too long for a GPU test to wait on.

    python tests/golden/make_config_fixtures.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "block-dda_py_amd"))

from oracle import gre, scatterer as osc  # noqa: E402
from bdda_amd import workloads  # noqa: E402

L_FIXTURE = 8


def oracle_geometry(wl, fields):
    g = gre.GREShape(wl["r_v_base"], wl["bc_ratio"], wl["ab_ratio"], wl["beta"])
    pts = g.surface_from_field(fields[wl["name"] + "__s"])
    n, grid = g.cuboid_lattice(pts)
    is_in = gre.inside_mask(g.lattice_lf, g.distance_factor, n, gre.nearest_distance(grid, pts))
    return n, g.lattice_lf, grid, is_in


def main():
    with np.load(os.path.join(HERE, "gre_fields.npz")) as d:
        fields = {k: d[k] for k in d.files}
    wl = workloads.FULL_64
    n, lf, grid, is_in = oracle_geometry(wl, fields)
    m_p = np.array(wl["m_p_xyz"], dtype=np.complex128)
    tg = osc.Target(n, lf, grid, is_in, m_p)
    wl0 = workloads.wavelength(wl, lf, tg.N)
    euler = workloads.euler_angles(wl["L"], wl["orient_seed"])[:L_FIXTURE]
    inc = osc.Incident(wl0, wl["m_m"], euler)
    t0 = time.time()
    r = osc.solve(tg, inc, "bicgstab", wl["tol"], wl["itermax"])
    print(f"lattice {tuple(int(v) for v in n)}, N = {tg.N}, {r['iter_fin'] + 1} iterations, "
          f"err {r['err_fin']:.3e}, {time.time() - t0:.1f} s")
    out = os.path.join(HERE, "config_full64_L8.npz")
    np.savez_compressed(out, lattice_n=np.asarray(n), N=tg.N, wl_0=wl0, euler=euler, iter_fin=r["iter_fin"],
                        err_fin=r["err_fin"], history=np.asarray(r["history"]), S_fw_theta=r["S_fw_theta"],
                        S_fw_phi=r["S_fw_phi"], S_bk=r["S_bk"], C_abs=r["C_abs"], C_ext=r["C_ext"])
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
