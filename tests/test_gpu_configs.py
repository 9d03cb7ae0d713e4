"""The workloads' own configurations on the GPU.

* The small configuration (24^3 GRE, beta = 0.1, m_p = 1.5 + 0.01i, L = 8,
  float64 block-COCG-rq to 1e-5) against the float64 oracle on the same
  geometry: same iteration count, S(0) and S_bk within 1e-6.
* The 96^3 configuration (the same physics at four times the size, L = 32)
  is too large for the oracle, so it is checked through what does not
  depend on size: the mixed-precision solve reaches the float64 true
  residual and its S(0) / S_bk agree with the float64 solve within 1e-4.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from oracle import scatterer as osc  # noqa: E402


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_small_configuration_matches_oracle():
    from bdda_amd import workloads
    from bl_dda.scatterer import DiscreteDipoles
    DiscreteDipoles.verbose_default = False
    wl = workloads.SMALL_24
    say(f"small configuration {wl['name']}")
    geo = workloads.build_geometry(wl)
    dd = workloads.make_dipoles(wl, geometry=geo, precision="fp64")
    assert max(dd.lattice_n) == 24
    dd.solve_matrix_equation()
    name, n, lf, grid, is_in = geo
    tg = osc.Target(n, lf, grid, is_in, np.array(wl["m_p_xyz"], dtype=np.complex128))
    inc = osc.Incident(dd.wl_0, wl["m_m"], workloads.euler_angles(wl["L"], wl["orient_seed"]))
    ref = osc.solve(tg, inc, "cocg", wl["tol"], wl["itermax"])
    assert dd.converge and dd.iter_fin == ref["iter_fin"]
    st, sp = dd.compute_PCAS_observable_S_fw()
    sb = dd.compute_OCBS_observable_S_bk()
    for a, b in [(st, ref["S_fw_theta"]), (sp, ref["S_fw_phi"]), (sb, ref["S_bk"])]:
        assert np.max(np.abs(a - b) / np.abs(b)) < 1e-6


def test_large_configuration_mixed_meets_fp64():
    from bdda_amd import workloads
    from bl_dda.scatterer import DiscreteDipoles
    DiscreteDipoles.verbose_default = False
    wl = workloads.LARGE_96
    say(f"large configuration {wl['name']}")
    dd = workloads.make_dipoles(wl, precision="fp64")
    assert dd._plan.n == (96, 96, 96)
    dd.solve_matrix_equation()
    assert dd.converge and dd.true_residual < 1e-5
    st64, sp64 = dd.compute_PCAS_observable_S_fw()
    sb64 = dd.compute_OCBS_observable_S_bk()
    dd.precision = "mixed"
    dd.solve_matrix_equation()
    assert dd.converge and dd.true_residual < 1e-5
    st, sp = dd.compute_PCAS_observable_S_fw()
    sb = dd.compute_OCBS_observable_S_bk()
    for a, b in [(st, st64), (sp, sp64), (sb, sb64)]:
        assert np.max(np.abs(a - b) / np.abs(b)) < 1e-4
