"""bdda_observables_noalloc (the allocation-free observables entry point)
against bdda_observables through the C ABI: two calls in a row on the same
GPU, the second on a smaller problem (fewer dipoles and orientations), each
equal to bdda_observables bit for bit; L beyond the fixed scratch size is
rejected with BDDA_E_ARG."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dipoles(r_v, L, seed):
    from bl_dda.scatterer import DiscreteDipoles, IncidentField, Target
    from oracle import gre, orient
    name, n, lf, grid, is_in, _, _ = gre.build_gre_geometry(r_v, 1.2, 1.1, 0.0, seed=seed)
    dd = DiscreteDipoles(Target(name, n, lf, grid, is_in, np.array([1.5 + 0.01j, 1.55 + 0.01j, 1.6 + 0.02j])),
                         IncidentField(0.5, 1.1, orient.euler_run_dda(L, seed=seed + 1)))
    dd.verbose = False
    dd.tol, dd.itermax = 1e-6, 80
    dd.set_interaction_matrix()
    dd.solve_matrix_equation()
    assert dd.converge
    return dd


def _inputs(dd):
    e0 = dd.e0_inc_vec_P
    geo = np.concatenate([dd.u_sca_fw_vec_P, dd.theta_sca_fw_vec_P, dd.phi_sca_fw_vec_P, dd.u_sca_bk_vec_P,
                          dd.theta_sca_bk_vec_P, dd.phi_sca_bk_vec_P,
                          np.stack([e0.real, e0.imag], axis=-1).reshape(dd.L, 6)], axis=1)
    epsfac = np.full(3, 0, dtype=np.complex128)
    epsfac[:] = (4 * np.pi) / ((dd.eper_r[0] - 1) * dd.element_vol)
    return np.ascontiguousarray(geo), epsfac


def _noalloc(dd, geo, epsfac, L=None):
    from bdda_amd import _lib
    lib = _lib.load()
    L = dd.L if L is None else L
    X = dd.X_device
    g = torch.from_numpy(geo).cuda()
    f = torch.from_numpy(epsfac).cuda()
    out = torch.full((max(L, 1), 5), complex("nan"), dtype=torch.complex128, device="cuda")
    s = torch.cuda.current_stream()
    rc = lib.bdda_observables_noalloc(dd._plan.h, ctypes.c_void_p(X.data_ptr()), L, ctypes.c_void_p(g.data_ptr()),
                                      ctypes.c_void_p(f.data_ptr()), ctypes.c_double(dd.k),
                                      ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(s.cuda_stream))
    s.synchronize()
    return rc, out.cpu().numpy()


def test_noalloc_observables_match_two_calls_in_a_row():
    big = _dipoles(0.09, 12, 3)
    small = _dipoles(0.06, 5, 4)
    assert small.num_element_occupy < big.num_element_occupy
    for dd in (big, small):
        say(f"noalloc observables N={dd.num_element_occupy} L={dd.L}")
        geo, epsfac = _inputs(dd)
        ref = dd._plan.observables(dd.X_device, geo, epsfac, dd.k)
        rc, got = _noalloc(dd, geo, epsfac)
        assert rc == 0
        assert np.array_equal(got, ref)
        # and the derived observables agree with the drop-in's own path
        st, sp = dd.compute_PCAS_observable_S_fw()
        assert np.array_equal(dd.k ** 2 * np.sqrt(2) * got[:, 0], st)


def test_noalloc_observables_reject_oversize_L():
    from bdda_amd import _lib
    dd = _dipoles(0.05, 3, 5)
    geo, epsfac = _inputs(dd)
    rc, _ = _noalloc(dd, geo, epsfac, L=257)
    assert rc == _lib.E_ARG
    assert b"256" in _lib.load().bdda_last_error()
