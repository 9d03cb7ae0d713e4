"""CPU-side checks of the product: the C-ABI library loads and exports every
entry point declared in include/bdda.h, and the host-side logic of the
drop-in modules matches the oracle (no GPU compute here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "bdda.h")
LIB = os.path.join(ROOT, "block-dda_py_amd", "lib", "libbdda.so")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bdda_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    for s in ["bdda_plan_create", "bdda_green_init", "bdda_mvp_offdiag", "bdda_block_mvp", "bdda_solve",
              "bdda_observables", "bdda_nearest_distance", "bdda_inside_mask", "bdda_last_error"]:
        assert s in syms


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbdda.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    for s in declared_symbols():
        assert hasattr(lib, s), s
    lib.bdda_version.restype = ctypes.c_int
    assert lib.bdda_version() == 100


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbdda.so not built")
def test_python_binding_covers_header():
    from bdda_amd import _lib
    assert sorted(_lib.exported_symbols()) == declared_symbols()
    lib = _lib.load()
    assert lib.bdda_last_error() == b""


def test_scatter_index_matches_oracle():
    from bl_krylov.bl_krylov import _build_scatter_index
    from oracle.bl_krylov import scatter_index
    addr = np.array([0, 5, 6, 17, 40])
    assert np.array_equal(_build_scatter_index(addr, 3), scatter_index(addr, 3))


def test_host_setup_matches_oracle():
    """Target / IncidentField host set-up equals the oracle's bit for bit."""
    from bl_dda import scatterer as prod
    from oracle import gre, orient, scatterer as ora
    name, n, lf, grid, is_in, rve, _ = gre.build_gre_geometry(0.08, 1.3, 1.1, 0.0)
    m_p = np.array([1.5 + 0.01j, 1.55 + 0.01j, 1.6 + 0.01j])
    tp = prod.Target(name, n, lf, grid, is_in, m_p)
    to = ora.Target(n, lf, grid, is_in, m_p)
    assert np.array_equal(tp.lattice_address_in_target, to.address)
    assert tp.ve_radius == to.ve_radius
    eu = orient.euler_run_dda(7, seed=3)
    ip = prod.IncidentField(0.5, 1.2, eu)
    io = ora.Incident(0.5, 1.2, eu)
    for a, b in [(ip.u_inc_vec_P, io.u_inc), (ip.e0_inc_vec_P, io.e0), (ip.theta_sca_bk_vec_P, io.th_bk),
                 (ip.phi_sca_fw_vec_P, io.ph_fw), (ip.u_sca_bk_vec_P, io.u_bk)]:
        assert np.array_equal(a, b)


def test_surface_points_match_oracle():
    from oracle import gre
    from shape_model.gaussian_ellipsoid import gaussian_ellipsoid_shape_model as G
    g = G(0.2, 1.4, 1.2, 0.15)
    o = gre.GREShape(0.2, 1.4, 1.2, 0.15)
    p1, _ = g.compute_r_points_on_GRE(np.random.default_rng(5))
    p2 = o.surface_points(np.random.default_rng(5))
    assert np.array_equal(p1, p2)
    d1, n1, g1 = g.create_cuboid_lattice_that_encloses_GRE_shape(p1)
    n2, g2 = o.cuboid_lattice(p2)
    assert np.array_equal(n1, n2) and np.array_equal(g1, g2)
    assert g.lattice_lf == o.lattice_lf and g.distance_factor == o.distance_factor


def test_batch_entry_validates_arguments_without_a_gpu():
    """bdda_solve_batch rejects an empty batch and a plan shared by two
    instances before touching the device."""
    import ctypes
    from bdda_amd import _lib
    lib = _lib.load()
    vp = ctypes.c_void_p
    nul = ctypes.POINTER(ctypes.c_int)()
    nd = ctypes.POINTER(ctypes.c_double)()
    rc = lib.bdda_solve_batch(0, None, 0, 0, None, None, None, None, 1e-5, 10, 20, 2, nul, nd, nd, nd, nul)
    assert rc == 1 and b"nb >= 1" in lib.bdda_last_error()
    fake = (vp * 2)(0x1000, 0x1000)
    bufs = (vp * 2)(0x2000, 0x3000)
    Ls = (ctypes.c_int * 2)(4, 4)
    rc = lib.bdda_solve_batch(2, fake, 0, 0, bufs, None, bufs, Ls, 1e-5, 10, 20, 2, nul, nd, nd, nd, nul)
    assert rc == 1 and b"own" in lib.bdda_last_error()


def test_plan_and_gemm_entries_validate_arguments_without_a_gpu():
    """The per-plan settings refuse a null plan; the GEMM entry points refuse
    an unknown epilogue and return at once on an empty batch, all before
    touching the device."""
    import ctypes
    from bdda_amd import _lib
    lib = _lib.load()
    assert lib.bdda_plan_set_x_kernel(None, 1) == 1 and b"null plan" in lib.bdda_last_error()
    assert lib.bdda_plan_x_kernel(None) == -1
    assert lib.bdda_plan_set_column_tile(None, 4) == 1
    fake = ctypes.c_void_p(0x1000)
    args = (1, 4, 4, 4, ctypes.c_float(1.0), fake, 0, 4, 1, fake, 0, 4, 1, ctypes.c_float(0.0), fake, 0, 4, 1)
    assert lib.bdda_gemm_bf16_ex(*args, None, 2, None) == 1 and b"bad arguments" in lib.bdda_last_error()
    empty = (0,) + args[1:]
    assert lib.bdda_gemm_bf16(*empty, None) == 0
    assert lib.bdda_gemm_bf16_ex(*empty, None, 1, None) == 0
