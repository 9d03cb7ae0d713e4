"""Batched uneven instances (bdda_solve_batch) against one-at-a-time solves.

Every instance keeps its own plan, so a batch must return, instance by
instance, exactly what DiscreteDipoles.solve_matrix_equation returns (same
kernels, same reduction order): X bit-identical, same iteration count and
residual history, for lattices of different sizes (register and
runtime-radix FFT lengths) in float64 and mixed precision.
"""
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


@pytest.fixture(scope="module")
def small_batch():
    from bdda_amd import workloads
    from bl_dda.scatterer import DiscreteDipoles
    DiscreteDipoles.verbose_default = False
    wl = dict(workloads.BATCH_32, grid=(7, 15), L=5, itermax=200)
    insts = workloads.batch_instances(wl, count=5)
    return [(w, workloads.build_geometry(w)) for w in insts]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("la_path", [0, 1])
def test_batch_matches_single_solves(small_batch, precision, la_path):
    """Host (0) or device (1) L x L algebra: the batch equals one-at-a-time
    solves bit for bit either way (each instance's device algebra works in
    its own plan's buffers on its own stream)."""
    from bdda_amd import _lib, workloads
    from bdda_amd.batch import solve_batch
    say(f"batch of {len(small_batch)} uneven instances, {precision}, la path {la_path}")
    lib = _lib.load()
    _lib.check(lib.bdda_set_la_path(la_path))
    try:
        dds = [workloads.make_dipoles(w, geometry=g, precision=precision) for w, g in small_batch]
        assert len({tuple(dd._plan.m) for dd in dds}) > 1  # uneven FFT grids
        ref = []
        for dd in dds:
            dd.solve_matrix_equation()
            assert dd.converge and dd.true_residual < 1e-5
            ref.append((dd.X_device.clone(), dd.iter_fin, list(dd.err_history)))
        solve_batch(dds, nstreams=3)
    finally:
        lib.bdda_set_la_path(0)
    for dd, (X, it, hist) in zip(dds, ref):
        assert torch.equal(dd.X_device, X)
        assert dd.iter_fin == it and dd.err_history == hist and dd.converge


def test_batch_rejects_shared_plans(small_batch):
    from bdda_amd import workloads
    from bdda_amd.batch import solve_batch
    w, g = small_batch[0]
    dd = workloads.make_dipoles(w, geometry=g, precision="fp64")
    with pytest.raises(ValueError, match="own"):
        solve_batch([dd, dd])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotuned_tiles_keep_the_operator(small_batch, precision):
    """bdda_plan_autotune fixes one of the candidate column tiles and x pass
    kernels; the operator at that setting equals the automatic one's to
    rounding, and the batch still converges to the tolerance with the tuned
    plans."""
    from bdda_amd import workloads
    from bdda_amd.batch import autotune, solve_batch
    say(f"autotuned column tiles, {precision}")
    dds = [workloads.make_dipoles(w, geometry=g, precision=precision) for w, g in small_batch]
    g = torch.Generator(device="cuda").manual_seed(3)
    for dd in dds:
        plan, L = dd._plan, dd._B_dev.shape[1]
        auto = plan.column_tile(L)
        S = torch.randn(dd._B_dev.shape, dtype=torch.complex128, device="cuda", generator=g)
        ref = plan.block_mvp(S)
        lb, xk = plan.autotune(L, precision)
        assert lb in {auto, 4, 8} and plan.column_tile(L) == lb and xk in {1, 2} and plan.x_kernel() == xk
        got = plan.block_mvp(S)
        assert float((got - ref).abs().max() / ref.abs().max()) < 1e-13
        plan.set_column_tile(0)
        plan.set_x_kernel(0)
        assert plan.column_tile(L) == auto and plan.x_kernel() == 0
    picks = autotune(dds)
    assert all(dd._plan.column_tile(dd._B_dev.shape[1]) == lb and dd._plan.x_kernel() == xk
               for dd, (lb, xk) in zip(dds, picks))
    solve_batch(dds, nstreams=3)
    assert all(dd.converge and dd.true_residual < 1e-5 for dd in dds)


def test_column_tile_out_of_range_is_refused(small_batch):
    from bdda_amd import workloads
    w, g = small_batch[0]
    dd = workloads.make_dipoles(w, geometry=g, precision="fp64")
    with pytest.raises(ValueError, match="tile"):
        dd._plan.set_column_tile(17)
    with pytest.raises(ValueError, match="tile"):
        dd._plan.set_column_tile(-1)


@pytest.mark.parametrize("draw", range(3))
def test_random_batches_match_single_solves(draw):
    """Seeded random batches: six particles of 6-14 cells, each with its own
    L of 1-40 (prime widths above the tile cap take 1-column tiles), a
    random precision, L x L path and stream count.  Each instance must come
    out of the batch bit-identical to its one-at-a-time solve."""
    from bdda_amd import _lib, workloads
    from bdda_amd.batch import solve_batch
    rng = np.random.default_rng(3000 + draw)
    precision, la_path = str(rng.choice(["fp64", "mixed"])), int(rng.integers(0, 2))
    nstreams = int(rng.integers(1, 7))
    wl = dict(workloads.BATCH_32, grid=(6, 14), itermax=300)
    insts = workloads.batch_instances(wl, count=12)
    pick = rng.choice(len(insts), size=6, replace=False)
    insts = [dict(insts[i], L=int(rng.integers(1, 41))) for i in pick]
    say(f"random batch draw={draw} L={[w['L'] for w in insts]} {precision} la={la_path} streams={nstreams}")
    lib = _lib.load()
    _lib.check(lib.bdda_set_la_path(la_path))
    try:
        dds = [workloads.make_dipoles(w, geometry=workloads.build_geometry(w), precision=precision) for w in insts]
        ref = []
        for dd in dds:
            dd.solve_matrix_equation()
            ref.append((dd.X_device.clone(), dd.iter_fin, list(dd.err_history), dd.converge))
        solve_batch(dds, nstreams=nstreams)
    finally:
        lib.bdda_set_la_path(0)
    for dd, (X, it, hist, conv) in zip(dds, ref):
        assert torch.equal(dd.X_device, X)
        assert dd.iter_fin == it and dd.err_history == hist and dd.converge == conv
