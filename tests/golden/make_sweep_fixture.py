"""Write tests/golden/sweep_oracle.npz: the float64 oracle's results for the
16 conditions of the reference's stored sweep
(``dda_results/pcas_ocbs_simulated_data.hdf5``, extracted into
``pcas_ocbs_simulated_data.npz``): beta = 0 oblate GREs (bc 3, r_v 0.1-0.4
um, geometry seed 12345), 2 (wavelength, medium) pairs x 2 m_p, 100
orientations from seed 12346, block-BiCGSTAB with the reference's defaults
(tol 1e-2, itermax 25; ``run_dda.py:38-70``, ``scatterer.py:115-116``).
Per condition: iteration count, C_abs, C_ext, S(0)_theta, S(0)_phi, S_bk,
and how far each moves when the orientations are perturbed by 1e-13 rad
(``sensitivity__*``): with 100 orientations of a particle of a few thousand
dipoles the L x L systems of the first iterations are ill-conditioned, so
rounding alone moves a 1-2 iteration result by up to ~1e-4 on the smallest
particle; a GPU result is held to a bound scaled from that.

    python tests/golden/make_sweep_fixture.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import gre, orient, scatterer as osc  # noqa: E402


def main():
    with np.load(os.path.join(HERE, "pcas_ocbs_simulated_data.npz")) as d:
        h5 = {k: d[k] for k in d.files}
    pairs = h5["target__wl_m_m_pairs"]
    mps = h5["target__m_p_xyz_list"]
    rvs = h5["target__r_v_base_list"]
    bc, ab, bt = (float(h5["target__" + k][0]) for k in ("bc_ratio_list", "ab_ratio_list", "gre_beta_list"))
    L = h5["target__simulated_data__C_abs"].shape[-1]
    euler = orient.euler_run_dda(L, seed=12346)
    shape = (len(pairs), len(mps), len(rvs), 1, 1, 1)
    out = {k: np.zeros(shape + (L,), dtype=dt) for k, dt in
           [("C_abs", np.float64), ("C_ext", np.float64), ("S_fw_PCAS_theta", np.complex128),
            ("S_fw_PCAS_phi", np.complex128), ("S_bk_OCBS", np.complex128)]}
    out["iter_fin"] = np.zeros(shape, dtype=np.int64)
    for k in list(out):
        if k != "iter_fin":
            out["sensitivity__" + k] = np.zeros(shape)
    out["r_ve"] = np.zeros((len(rvs), 1, 1, 1))
    for i_rv, rv in enumerate(rvs):
        _, n, lf, grid, is_in, r_ve, _ = gre.build_gre_geometry(rv, bc, ab, bt, seed=12345)
        out["r_ve"][i_rv] = r_ve
        for i_p, (wl, mm) in enumerate(pairs):
            for i_mp, m_p in enumerate(mps):
                t0 = time.time()
                tg = osc.Target(n, lf, grid, is_in, np.asarray(m_p))
                r = osc.solve(tg, osc.Incident(wl, mm, euler), "bicgstab", 1e-2, 25)
                assert r["converge"]
                jitter = np.random.default_rng(7).uniform(-1e-13, 1e-13, euler.shape)
                rp = osc.solve(tg, osc.Incident(wl, mm, euler + jitter), "bicgstab", 1e-2, 25)
                idx = (i_p, i_mp, i_rv, 0, 0, 0)
                out["iter_fin"][idx] = r["iter_fin"]
                for k, rk in [("C_abs", "C_abs"), ("C_ext", "C_ext"), ("S_fw_PCAS_theta", "S_fw_theta"),
                              ("S_fw_PCAS_phi", "S_fw_phi"), ("S_bk_OCBS", "S_bk")]:
                    out[k][idx] = r[rk]
                    den = np.max(np.abs(r[rk]))
                    out["sensitivity__" + k][idx] = np.max(np.abs(rp[rk] - r[rk])) / den if den > 0 else 0.0
                print(f"r_v {rv} lattice {tuple(int(v) for v in n)} pair {i_p} m_p {i_mp}: "
                      f"{r['iter_fin'] + 1} iterations, {time.time() - t0:.1f} s", flush=True)
    path = os.path.join(HERE, "sweep_oracle.npz")
    np.savez_compressed(path, euler=euler, **out)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
