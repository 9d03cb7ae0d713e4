"""Pins the CPU oracle against the reference's recorded data and against
what the mathematics fixes (closed forms, brute force, dense solves)."""
import numpy as np
import pytest

from oracle import bl_krylov, dense, gre, mie, mvp_fft, orient, scatterer

SD = "target__simulated_data__"


# ----------------------------------------------------------------- GRE shape
def test_gre_notebook_lattice(golden_nb):
    nb = golden_nb
    name, n, lf, grid, is_in, rve, _ = gre.build_gre_geometry(
        float(nb["r_v_base"]), 1, 1, 0.0, seed=int(nb["shape_seed"]))
    assert n.tolist() == nb["lattice_n"].tolist()
    m_p = nb["m_p_xyz"]
    dpl = (float(nb["wl_0"]) / np.abs(np.max(m_p))) / lf
    assert dpl == float(nb["dpl"])  # recorded at 16 digits: bit-exact
    # the Mie Q values recorded at full precision fix r_ve, hence N_occ
    tg = scatterer.Target(n, lf, grid, is_in, m_p)
    q = mie.mie_q_and_s(float(nb["wl_0"]), float(nb["m_m"]), tg.ve_radius, 1.5 + 0.01j)
    np.testing.assert_allclose(q[:3], nb["mie_q"], rtol=1e-13)


@pytest.mark.parametrize("irv", [0, 1, 2, 3])
def test_gre_hdf5_r_ve_bit_exact(golden_h5, irv):
    rv = golden_h5["target__r_v_base_list"][irv]
    bc = golden_h5["target__bc_ratio_list"][0]
    ab = golden_h5["target__ab_ratio_list"][0]
    bt = golden_h5["target__gre_beta_list"][0]
    name, n, lf, grid, is_in, rve, _ = gre.build_gre_geometry(rv, bc, ab, bt)
    stored = golden_h5[SD + "r_ve"][irv, 0, 0, 0]
    # the occupied count the stored r_ve encodes is an exact integer match
    assert int(round(4 * np.pi * stored ** 3 / (3 * lf ** 3))) == int(is_in.sum())
    # r_ve itself agrees to 2 ulp (libm cbrt/pow differ between NumPy builds)
    assert abs(rve - stored) <= 2 * np.spacing(stored)


def test_nearest_distance_brute_equals_kdtree():
    g = gre.GREShape(0.15, 1.3, 1.2, 0.1)
    pts = g.surface_points(np.random.default_rng(3))
    n, grid = g.cuboid_lattice(pts)
    sel = grid[:: max(1, len(grid) // 3000)]
    a = gre.nearest_distance(sel, pts)
    b = gre.nearest_distance_brute(sel, pts)
    assert np.array_equal(a, b)


def test_inside_mask_sphere_closed_form():
    # beta = 0 sphere of radius R: every lattice point deeper than the
    # near-surface shell is inside, and no inside point lies outside the sphere
    R = 0.15
    g = gre.GREShape(R, 1, 1, 0.0)
    name, n, lf, grid, is_in, rve, _ = gre.build_gre_geometry(R, 1, 1, 0.0, seed=0)
    r = np.linalg.norm(grid, axis=1)
    deep = r < R - 1.01 * g.distance_factor * lf
    assert deep.sum() > 100 and np.all(is_in[deep])
    assert np.all(r[is_in] < R + g.distance_factor * lf)
    assert not np.any(is_in[r > R + g.distance_factor * lf])


# ------------------------------------------------------------- orientations
def test_euler_draw_matches_hdf5(golden_h5):
    e = golden_h5[SD + "Euler_angles"]
    ref = orient.euler_run_dda(100)
    assert np.array_equal(e[0, 0, 0, 0, 0, 0], ref)
    assert np.array_equal(e[1, 1, 3, 0, 0, 0], ref)


def test_euler_draw_matches_notebook(golden_nb):
    np.testing.assert_allclose(np.degrees(orient.euler_test_dda()), golden_nb["euler_deg"], atol=5e-9)


def test_rotation_closed_form():
    eu = np.array([[0.3, 0.7, 1.1], [2.0, 0.2, -0.5]])
    inc = scatterer.Incident(0.5, 1.0, eu)

    def rz(t):
        c, s = np.cos(t), np.sin(t)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])

    def ry(t):
        c, s = np.cos(t), np.sin(t)
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])

    for l, (a, b, g) in enumerate(eu):
        R = (rz(a) @ ry(b) @ rz(g)).T  # lab -> particle
        np.testing.assert_allclose(inc.rot[l], R, atol=1e-14)


# -------------------------------------------------------------- operator
def test_green_tensor_closed_form_and_symmetry():
    n, lf, k = (3, 4, 5), 0.07, 9.3
    A = mvp_fft.interaction_tensor(n, lf, k)
    assert A.shape == (6, 8, 10, 3, 3)
    assert np.array_equal(A, np.swapaxes(A, -1, -2))  # exact symmetry
    assert np.all(A[0, 0, 0] == 0) and np.all(A[3] == 0) and np.all(A[:, 4] == 0) and np.all(A[:, :, 5] == 0)
    for I in [(1, 2, 3), (5, 7, 9), (2, 0, 6)]:
        d = [(I[a] if I[a] <= n[a] else I[a] - 2 * n[a]) * lf for a in range(3)]
        np.testing.assert_allclose(A[I], mvp_fft.green_element(*d, k), rtol=1e-13)


@pytest.mark.parametrize("n", [(3, 3, 3), (4, 3, 5), (2, 5, 3)])
def test_fft_mvp_equals_dense_sum(n):
    lf, k = 0.05, 11.0
    rng = np.random.default_rng(1)
    ncell = int(np.prod(n))
    P = rng.standard_normal((3 * ncell, 4)) + 1j * rng.standard_normal((3 * ncell, 4))
    At = mvp_fft.fft_init(n, 3, lf, k)
    Y = mvp_fft.fft_mvp(n, 3, At, P)
    D = dense.dense_offdiag(n, lf, k)
    np.testing.assert_allclose(Y, D @ P, rtol=0, atol=1e-11 * np.abs(D @ P).max())
    y1 = mvp_fft.fft_mvp(n, 3, At, P[:, 0])
    np.testing.assert_allclose(y1, Y[:, 0], rtol=1e-14, atol=1e-14 * np.abs(y1).max())


def test_fft_init_rejects_f():
    with pytest.raises(ValueError):
        mvp_fft.fft_init((2, 2, 2), 2, 0.1, 1.0)


def test_operator_complex_symmetric():
    n, lf, k = (4, 4, 3), 0.06, 10.0
    rng = np.random.default_rng(2)
    At = mvp_fft.fft_init(n, 3, lf, k)
    x = rng.standard_normal(3 * 48) + 1j * rng.standard_normal(3 * 48)
    y = rng.standard_normal(3 * 48) + 1j * rng.standard_normal(3 * 48)
    a = x @ mvp_fft.fft_mvp(n, 3, At, y)
    b = y @ mvp_fft.fft_mvp(n, 3, At, x)
    assert abs(a - b) < 1e-12 * abs(a)


# --------------------------------------------------------------- solvers
def _small_problem(m_p=(1.5 + 0.01j,) * 3, L=5, seed=0):
    name, n, lf, grid, is_in, rve, _ = gre.build_gre_geometry(0.06, 1.2, 1.1, 0.0, seed=seed)
    tg = scatterer.Target(n, lf, grid, is_in, np.array(m_p))
    inc = scatterer.Incident(0.3, 1.2, orient.euler_run_dda(L, seed=seed + 1))
    return tg, inc


@pytest.mark.parametrize("solver", ["cocg", "bicgstab"])
def test_solvers_converge_to_dense_solution(solver):
    tg, inc = _small_problem()
    assert 30 < tg.N < 400
    eps_r, alpha, diag = scatterer.polarizability(tg, inc)
    At = mvp_fft.fft_init(tg.lattice_n, 3, tg.lf, inc.k)
    E_inc, B = scatterer.rhs(tg, inc)
    A = dense.dense_offdiag(tg.lattice_n, tg.lf, inc.k, cells=tg.address) + np.diag(diag)
    Xd = np.linalg.solve(A, B)
    fn = bl_krylov.bl_cocg_rq if solver == "cocg" else bl_krylov.bl_bicgstab
    X, it, err, hist = fn(tg.lattice_n, 3, tg.address, At, diag, B, 1e-10, 200)
    assert err < 1e-10 and it == len(hist) - 1
    np.testing.assert_allclose(X, Xd, rtol=0, atol=1e-8 * np.abs(Xd).max())


def test_cocg_invariant_to_qr_basis():
    # Q -> Q D, xi -> D* xi for a diagonal unitary D leaves X unchanged
    tg, inc = _small_problem(L=4)
    eps_r, alpha, diag = scatterer.polarizability(tg, inc)
    At = mvp_fft.fft_init(tg.lattice_n, 3, tg.lf, inc.k)
    E_inc, B = scatterer.rhs(tg, inc)
    X1, *_ = bl_krylov.bl_cocg_rq(tg.lattice_n, 3, tg.address, At, diag, B, 1e-12, 6)
    orig = np.linalg.qr
    D = np.exp(1j * np.array([0.3, -1.2, 2.0, 0.7]))

    def qr_phase(a, mode='reduced'):
        q, r = orig(a, mode=mode)
        return q * D[None, :], np.conj(D)[:, None] * r

    np.linalg.qr = qr_phase
    try:
        X2, *_ = bl_krylov.bl_cocg_rq(tg.lattice_n, 3, tg.address, At, diag, B, 1e-12, 6)
    finally:
        np.linalg.qr = orig
    np.testing.assert_allclose(X2, X1, rtol=0, atol=1e-12 * np.abs(X1).max())


def test_notebook_end_to_end(golden_nb):
    nb = golden_nb
    name, n, lf, grid, is_in, rve, _ = gre.build_gre_geometry(
        float(nb["r_v_base"]), 1, 1, 0.0, seed=int(nb["shape_seed"]))
    tg = scatterer.Target(n, lf, grid, is_in, nb["m_p_xyz"])
    inc = scatterer.Incident(float(nb["wl_0"]), float(nb["m_m"]), orient.euler_test_dda())
    out = scatterer.solve(tg, inc, "bicgstab", 1e-2, 25)
    assert out["iter_fin"] == int(nb["iter_fin"])
    assert ["%.4f" % e for e in out["history"]] == ["%.4f" % e for e in nb["err_printed"]]
    # S_bk recorded to 8 significant digits
    np.testing.assert_allclose(out["S_bk"], nb["s_bk"], rtol=2e-6, atol=1e-8)
    q = mie.mie_q_and_s(float(nb["wl_0"]), float(nb["m_m"]), tg.ve_radius, 1.5 + 0.01j)
    qabs = out["C_abs"] / (np.pi * tg.ve_radius ** 2)
    qext = out["C_ext"] / (np.pi * tg.ve_radius ** 2)
    np.testing.assert_allclose(np.abs(qabs / q[1] - 1), nb["qabs_relerr"], rtol=5e-5, atol=1e-8)
    np.testing.assert_allclose(np.abs(qext / q[2] - 1), nb["qext_relerr"], rtol=5e-5, atol=1e-8)
    assert q[3] == nb["mie_s_fw"] and q[4] == nb["mie_s_bk"]


def test_hdf5_conditions_weak(golden_h5):
    """The stored sweep results (r_v = 0.1 um, 4 conditions) agree with the
    restated path to the solver tolerance class (1e-2); the stored run's
    iterates are not reproduced exactly by reference v0.3.0 (see DESIGN)."""
    g = golden_h5
    name, n, lf, grid, is_in, rve, _ = gre.build_gre_geometry(0.1, 3.0, 1.0, 0.0)
    for ip in range(2):
        for im in range(2):
            eu = g[SD + "Euler_angles"][ip, im, 0, 0, 0, 0][:20]
            tg = scatterer.Target(n, lf, grid, is_in, g["target__m_p_xyz_list"][im])
            wl, mm = g["target__wl_m_m_pairs"][ip]
            out = scatterer.solve(tg, scatterer.Incident(wl, mm, eu), "bicgstab", 1e-2, 25)
            for key, ds in [("S_fw_theta", "S_fw_PCAS_theta"), ("S_fw_phi", "S_fw_PCAS_phi"),
                            ("S_bk", "S_bk_OCBS"), ("C_ext", "C_ext")]:
                ref = g[SD + ds][ip, im, 0, 0, 0, 0][:20]
                assert np.max(np.abs(out[key] - ref) / np.abs(ref)) < 2e-2


# -------------------------------------------------------------- physics
def test_single_dipole_closed_form():
    grid = np.zeros((1, 3))
    tg = scatterer.Target(np.array([1, 1, 1]), 0.02, grid, np.array([True]), np.array([1.6 + 0.1j] * 3))
    inc = scatterer.Incident(0.5, 1.1, orient.euler_run_dda(3, seed=4))
    eps_r, alpha, diag = scatterer.polarizability(tg, inc)
    out = scatterer.solve(tg, inc, "bicgstab", 1e-12, 5)
    E_inc, B = scatterer.rhs(tg, inc)
    np.testing.assert_allclose(out["P"], alpha[None] * E_inc, rtol=1e-14)
    # S_fw = k^2 sqrt2 (P.th conj(e) ...): for one dipole at the origin
    a = alpha[0, 0]
    np.testing.assert_allclose(out["S_fw_theta"], inc.k ** 2 * a * np.sqrt(2) / np.sqrt(2) * np.ones(3), rtol=1e-13)


def test_mie_rayleigh_limit_and_optical_theorem():
    wl, mm, r, mp = 10.0, 1.0, 0.05, 1.5 + 0.1j
    Qs, Qa, Qe, Sf, Sb = mie.mie_q_and_s(wl, mm, r, mp)
    x = 2 * np.pi * r / wl
    m2 = mp ** 2
    Kc = (m2 - 1) / (m2 + 2)
    np.testing.assert_allclose(Qa, 4 * x * Kc.imag, rtol=2e-3)
    np.testing.assert_allclose(Qs, 8 / 3 * x ** 4 * abs(Kc) ** 2, rtol=2e-2)
    # optical theorem in the S11 = S2/(-ik) convention: C_ext = (4 pi / k) Im S_fw
    wl, mm, r = 0.6, 1.33, 0.3
    Qs2, Qa2, Qe2, Sf2, Sb2 = mie.mie_q_and_s(wl, mm, r, 1.5 + 0.01j)
    k2 = 2 * np.pi * mm / wl
    np.testing.assert_allclose(Qe2 * np.pi * r ** 2, 4 * np.pi / k2 * float(np.imag(Sf2)), rtol=1e-6)


# ------------------------------------------------------- learned warm start
def test_fno_oracle_spectral_equals_direct_dft():
    from oracle import fno
    G, K, C = 4, 2, 2
    rng = np.random.default_rng(0)
    v = rng.standard_normal((1, C, G ** 3))
    Rr = rng.standard_normal((K ** 3, C, C))
    Ri = rng.standard_normal((K ** 3, C, C))
    ks = fno.mode_indices(G, K)
    n = np.arange(G)
    idx = np.stack(np.meshgrid(n, n, n, indexing="ij"), -1).reshape(-1, 3)
    out = np.zeros((1, C, G ** 3))
    R = (Rr + 1j * Ri).reshape(K, K, K, C, C)
    for a, kx in enumerate(ks):
        for b, ky in enumerate(ks):
            for c, kz in enumerate(ks):
                ph = np.exp(-2j * np.pi * (idx @ np.array([kx, ky, kz])) / G)
                vh = v[0] @ ph                       # (C,)
                yh = vh @ R[a, b, c]                 # (C,)
                out[0] += (yh[:, None] * np.conj(ph)[None, :]).real / G ** 3
    np.testing.assert_allclose(fno.spectral(v, Rr, Ri, G, K), out, rtol=1e-12, atol=1e-12)
