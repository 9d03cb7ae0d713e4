"""GPU parity of the HIP path (through the C ABI) against the float64 oracle.

Tolerances: float64 path 1e-12 (operator) / 1e-8 (solver iterates, which
amplify rounding by the iteration count) relative; S(0) and S_bk within 1e-6
relative in float64 mode and 1e-4 in mixed mode (the required bars); the GRE
voxel mask and distances bit-exact.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from oracle import bl_krylov as okr  # noqa: E402
from oracle import gre as ogre  # noqa: E402
from oracle import mvp_fft as omvp  # noqa: E402
from oracle import orient, scatterer as osc  # noqa: E402


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# --------------------------------------------------------------- operator
@pytest.mark.parametrize("n,L", [((4, 5, 6), 3), ((7, 5, 11), 9), ((8, 8, 8), 16), ((3, 9, 4), 1)])
def test_fft_mvp_matches_oracle(n, L):
    say(f"fft_mvp n={n} L={L}")
    from mvp_fft import mvp_fft as pm
    lf, k = 0.031, 12.7
    rng = np.random.default_rng(0)
    ncell = int(np.prod(n))
    P = rng.standard_normal((3 * ncell, L)) + 1j * rng.standard_normal((3 * ncell, L))
    ref = omvp.fft_mvp(n, 3, omvp.fft_init(n, 3, lf, k), P)
    At = pm.fft_init(n, 3, lf, k)
    Y = pm.fft_mvp(n, 3, At, P)
    assert Y.shape == ref.shape and Y.dtype == np.complex128
    assert rel(Y, ref) < 1e-12
    Ym = pm.fft_mvp(n, 3, At, P, precision="mixed")
    assert rel(Ym, ref) < 2e-6
    y1 = pm.fft_mvp(n, 3, At, P[:, 0])
    assert y1.shape == (3 * ncell,) and rel(y1, ref[:, 0]) < 1e-12


@pytest.mark.parametrize("n", [(8, 12, 20), (16, 24, 32), (40, 64, 3), (128, 2, 3), (10, 15, 18), (25, 23, 27),
                               (30, 36, 38), (45, 4, 41), (48, 5, 50), (54, 60, 3), (72, 4, 68), (75, 80, 2),
                               (90, 3, 4), (2, 113, 3), (4, 120, 3), (96, 4, 3)])
def test_register_and_stockham_kernels_agree(n):
    """Every register-kernel length (16 ... 256: radix pairs from
    {3,4,5,6,8,9,10,12,15,16}) on some axis, at a runtime column tile,
    against the runtime-radix kernel and the oracle, both precisions."""
    say(f"kernel paths n={n}")
    from bdda_amd import _lib
    from mvp_fft import mvp_fft as pm
    lib = _lib.load()
    lf, k = 0.029, 10.3
    rng = np.random.default_rng(1)
    ncell = int(np.prod(n))
    P = rng.standard_normal((3 * ncell, 5)) + 1j * rng.standard_normal((3 * ncell, 5))
    ref = omvp.fft_mvp(n, 3, omvp.fft_init(n, 3, lf, k), P)
    At = pm.fft_init(n, 3, lf, k)
    out = {}
    try:
        for path in (0, 1, 2):
            _lib.check(lib.bdda_set_fft_path(path))
            out[path] = (pm.fft_mvp(n, 3, At, P), pm.fft_mvp(n, 3, At, P, precision="mixed"))
    finally:
        lib.bdda_set_fft_path(0)
    for path in (0, 1, 2):
        assert rel(out[path][0], ref) < 1e-12
        assert rel(out[path][1], ref) < 2e-6


@pytest.mark.parametrize("n,L,lb", [((8, 5, 4), 16, 8), ((12, 5, 4), 16, 8), ((16, 5, 4), 16, 8),
                                    ((20, 5, 4), 16, 8), ((24, 5, 4), 16, 8), ((32, 4, 3), 16, 8),
                                    ((40, 5, 4), 16, 8), ((64, 6, 5), 16, 8), ((96, 4, 3), 16, 4),
                                    ((64, 6, 5), 4, 4), ((16, 5, 4), 4, 4), ((40, 5, 4), 4, 4),
                                    ((10, 5, 4), 16, 8), ((15, 5, 4), 16, 8), ((18, 5, 4), 16, 8),
                                    ((23, 5, 4), 16, 8), ((25, 5, 4), 16, 8), ((27, 5, 4), 16, 8),
                                    ((30, 5, 4), 16, 8), ((36, 5, 4), 16, 8), ((38, 5, 4), 16, 8),
                                    ((41, 5, 4), 16, 8), ((45, 5, 4), 16, 8), ((48, 5, 4), 16, 8),
                                    ((50, 5, 4), 16, 8), ((54, 5, 4), 16, 8), ((60, 5, 4), 16, 8),
                                    ((68, 5, 4), 16, 8), ((72, 5, 4), 16, 8), ((75, 5, 4), 16, 8),
                                    ((80, 5, 4), 16, 8), ((90, 4, 3), 16, 4), ((113, 4, 3), 16, 4),
                                    ((120, 4, 3), 16, 4), ((54, 5, 4), 4, 4), ((68, 5, 4), 4, 4)])
def test_x_pass_kernel_on_the_compiled_tile(n, L, lb):
    """The compiled 8- and 4-column tiles run the x pass in the dedicated
    kernel, which needs an even R1: the plan gives x a register length with
    one (2N, or 48, 80, 90, 144, 240 where the smallest 5-smooth length
    >= 2N - 1 would be 45, 75, 81, 135, 225: x = 23, 38, 41, 68, 113 here).
    It must agree with the generic register kernel (path 2), the
    runtime-radix kernels (path 1), the split-spectrum x pass (path 3) and
    the one-pass x pass (path 4) on every length, and the oracle, and repeat
    bit for bit."""
    say(f"x pass kernel n={n} L={L}")
    from bdda_amd import _lib
    from bdda_amd.plan import Plan
    from mvp_fft import mvp_fft as pm
    lib = _lib.load()
    plan = Plan(n)
    assert plan.column_tile(L) == lb
    assert plan.m[0] == {23: 48, 38: 80, 41: 90, 68: 144, 113: 240}.get(n[0], 2 * n[0])
    lf, k = 0.031, 8.7
    rng = np.random.default_rng(11)
    ncell = int(np.prod(n))
    P = rng.standard_normal((3 * ncell, L)) + 1j * rng.standard_normal((3 * ncell, L))
    ref = omvp.fft_mvp(n, 3, omvp.fft_init(n, 3, lf, k), P)
    At = pm.fft_init(n, 3, lf, k)
    out = {}
    try:
        for path in (0, 1, 2, 3, 4):
            _lib.check(lib.bdda_set_fft_path(path))
            out[path] = (pm.fft_mvp(n, 3, At, P), pm.fft_mvp(n, 3, At, P, precision="mixed"))
        _lib.check(lib.bdda_set_fft_path(0))
        again = (pm.fft_mvp(n, 3, At, P), pm.fft_mvp(n, 3, At, P, precision="mixed"))
    finally:
        lib.bdda_set_fft_path(0)
    for path in (0, 1, 2, 3, 4):
        assert rel(out[path][0], ref) < 1e-12, path
        assert rel(out[path][1], ref) < 2e-6, path
    assert rel(out[0][0], out[2][0]) < 1e-13
    # the split-spectrum x pass (path 3) and the one-pass one (path 4) on every length
    for path in (3, 4):
        assert rel(out[path][0], out[2][0]) < 1e-13 and rel(out[path][1], out[2][1]) < 1e-6, path
    # the dedicated kernel repeats bit for bit
    assert np.array_equal(out[0][0], again[0]) and np.array_equal(out[0][1], again[1])


@pytest.mark.parametrize("n,L", [((96, 4, 3), 8), ((4, 96, 3), 8), ((3, 4, 96), 8), ((3, 4, 96), 5)]
                         + [((3, 4, nz), 8) for nz in (10, 15, 18, 23, 25, 27, 30, 36, 38, 41, 45, 48, 50, 54, 60,
                                                       68, 72, 75, 80, 90, 113, 120)])
def test_block_mvp_lengths_match_oracle(n, L):
    """Length 192 (16 x 12: a 96-cell axis, as on the 96^3 cuboid) on each
    axis, and every other register length on z, through the solver's
    occupied-block path (scatter / gather z passes, Jacobi epilogue),
    against the oracle's _block_mvp; then the off-diagonal product on the
    cuboid in both precisions.  L = 8 takes a compiled tile, L = 5 a runtime
    one."""
    say(f"block mvp n={n} L={L}")
    from bl_krylov import bl_krylov as pk
    from mvp_fft import mvp_fft as pm
    lf, k = 0.03, 9.5
    rng = np.random.default_rng(7)
    ncell = int(np.prod(n))
    address = np.sort(rng.choice(ncell, size=ncell // 3, replace=False)).astype(np.int64)
    N = address.size
    diag = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N) + 3.0
    jpre = 1.0 / diag
    S = rng.standard_normal((3 * N, L)) + 1j * rng.standard_normal((3 * N, L))
    idx = pk._build_scatter_index(address, 3)
    a = pk._block_mvp(n, 3, pm.fft_init(n, 3, lf, k), idx, diag, S, jpre)
    b = okr.block_mvp(n, 3, omvp.fft_init(n, 3, lf, k), idx, diag, S, jpre)
    assert rel(a, b) < 1e-12
    # the same operator as the off-diagonal product on the cuboid, mixed
    P = rng.standard_normal((3 * ncell, L)) + 1j * rng.standard_normal((3 * ncell, L))
    ref = omvp.fft_mvp(n, 3, omvp.fft_init(n, 3, lf, k), P)
    At = pm.fft_init(n, 3, lf, k)
    assert rel(pm.fft_mvp(n, 3, At, P), ref) < 1e-12
    assert rel(pm.fft_mvp(n, 3, At, P, precision="mixed"), ref) < 2e-6


@pytest.mark.parametrize("n,L,box", [((20, 18, 16), 8, (7, 5, 16)), ((16, 24, 20), 8, (16, 9, 3)),
                                     ((12, 12, 12), 5, (1, 1, 1)), ((40, 6, 5), 16, (13, 6, 5)),
                                     ((64, 6, 5), 8, (44, 3, 5)), ((9, 10, 96), 8, (4, 10, 50))])
def test_block_mvp_on_a_confined_target_matches_oracle(n, L, box):
    """Every occupied cell inside the box at the cuboid's origin: the passes
    then skip the lines and rows beyond it (x rows, y rows, z lines), and the
    product must still equal the oracle's _block_mvp over the whole cuboid.
    The cases take the dedicated x pass (lengths 80 and 128, the latter in
    its register-lean form), the register z pass on length 192, and a
    runtime tile."""
    say(f"confined block mvp n={n} L={L} box={box}")
    from bl_krylov import bl_krylov as pk
    from mvp_fft import mvp_fft as pm
    from bdda_amd.plan import Plan
    lf, k = 0.03, 9.5
    rng = np.random.default_rng(17)
    bx, by, bz = box
    cells = np.array([(ix * n[1] + iy) * n[2] + iz for ix in range(bx) for iy in range(by) for iz in range(bz)])
    pick = rng.choice(cells.size, size=max(1, cells.size // 3), replace=False)
    corner = ((bx - 1) * n[1] + by - 1) * n[2] + bz - 1
    address = np.unique(np.append(cells[pick], corner)).astype(np.int64)
    N = address.size
    diag = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N) + 3.0
    jpre = 1.0 / diag
    S = rng.standard_normal((3 * N, L)) + 1j * rng.standard_normal((3 * N, L))
    idx = pk._build_scatter_index(address, 3)
    a = pk._block_mvp(n, 3, pm.fft_init(n, 3, lf, k), idx, diag, S, jpre)
    b = okr.block_mvp(n, 3, omvp.fft_init(n, 3, lf, k), idx, diag, S, jpre)
    # one occupied cell: the off-diagonal part is exactly zero (no other
    # dipole), so b = S and what the convolution leaves is its rounding,
    # eps |G(lf)| / |diag| ~ 1e-12 of S, whichever the radix order
    assert rel(a, b) < (1e-12 if N > 1 else 5e-12)
    pl = Plan(n)
    pl.set_target(address, diag, jpre)
    assert pl.occupied_extent == box


@pytest.mark.parametrize("draw", range(12))
def test_block_mvp_random_configurations_match_oracle(draw):
    """Seeded random configurations of the occupied-block operator: cuboid
    axes of 2-48 cells (any length the plan picks: register lengths, the
    runtime-radix kernels, balanced and unbalanced x radix pairs), L of
    1-70 (compiled 8- and 4-column tiles and runtime tiles), and the target
    a random fraction of a random sub-box anchored at the origin (so the
    passes prune to its extent) or scattered over the whole cuboid.
    Against the oracle's _block_mvp at 1e-12, and the off-diagonal cuboid
    product in complex64 mode at 2e-6."""
    from bl_krylov import bl_krylov as pk
    from mvp_fft import mvp_fft as pm
    rng = np.random.default_rng(1000 + draw)
    while True:
        n = tuple(int(v) for v in rng.integers(2, 49, size=3))
        if np.prod(n) <= 24000:
            break
    L = int(rng.integers(1, 71))
    confined = bool(rng.integers(0, 2))
    box = tuple(int(rng.integers(1, v + 1)) for v in n) if confined else n
    say(f"random block mvp draw={draw} n={n} L={L} box={box}")
    lf, k = 0.03, float(rng.uniform(5.0, 12.0))
    cells = np.array([(ix * n[1] + iy) * n[2] + iz for ix in range(box[0]) for iy in range(box[1])
                      for iz in range(box[2])])
    frac = float(rng.uniform(0.05, 1.0))
    address = np.unique(rng.choice(cells, size=max(1, int(frac * cells.size)), replace=False)).astype(np.int64)
    N = address.size
    diag = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N) + 3.0
    jpre = 1.0 / diag
    S = rng.standard_normal((3 * N, L)) + 1j * rng.standard_normal((3 * N, L))
    idx = pk._build_scatter_index(address, 3)
    a = pk._block_mvp(n, 3, pm.fft_init(n, 3, lf, k), idx, diag, S, jpre)
    b = okr.block_mvp(n, 3, omvp.fft_init(n, 3, lf, k), idx, diag, S, jpre)
    assert rel(a, b) < (1e-12 if N > 1 else 5e-12)
    ncell = int(np.prod(n))
    P = rng.standard_normal((3 * ncell, L)) + 1j * rng.standard_normal((3 * ncell, L))
    ref = omvp.fft_mvp(n, 3, omvp.fft_init(n, 3, lf, k), P)
    assert rel(pm.fft_mvp(n, 3, pm.fft_init(n, 3, lf, k), P, precision="mixed"), ref) < 2e-6


@pytest.mark.parametrize("draw", range(4))
def test_random_sequences_on_one_plan_match_oracle(draw):
    """One plan through a seeded random sequence of twelve steps: new
    targets (random sub-boxes at the origin or scattered cells), occupied
    block products and cuboid products (both precisions) of random widths,
    a Green spectrum rebuilt at another wavenumber, and column tiles forced
    or left automatic.  Every product must equal the oracle's for the
    plan's current state, so no cached line list, extent, zeroed scratch,
    workspace size or spectrum may outlive what it was built for."""
    from bdda_amd.plan import Plan
    rng = np.random.default_rng(4000 + draw)
    while True:
        n = tuple(int(v) for v in rng.integers(3, 21, size=3))
        if np.prod(n) <= 6000:
            break
    ncell = int(np.prod(n))
    lf, k = 0.03, float(rng.uniform(5.0, 12.0))
    p = Plan(n)
    p.green_init(lf, k)
    Ao = omvp.fft_init(n, 3, lf, k)
    tgt = None
    ops = []
    for step in range(12):
        op = str(rng.choice(["target", "block", "block", "cuboid", "green", "tile"])) if tgt else "target"
        if op == "target":
            box = tuple(int(rng.integers(1, v + 1)) for v in n) if rng.integers(0, 2) else n
            cells = np.array([(ix * n[1] + iy) * n[2] + iz for ix in range(box[0]) for iy in range(box[1])
                              for iz in range(box[2])])
            address = np.unique(rng.choice(cells, size=max(2, int(rng.uniform(0.1, 1.0) * cells.size)))).astype(
                np.int64)
            N = address.size
            diag = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N) + 3.0
            tgt = (address, diag, 1.0 / diag)
            p.set_target(*tgt)
            ops.append(f"target N={N} box={box}")
        elif op == "block":
            L = int(rng.integers(1, 41))
            address, diag, jpre = tgt
            S = rng.standard_normal((3 * address.size, L)) + 1j * rng.standard_normal((3 * address.size, L))
            idx = okr.scatter_index(address, 3)
            got = p.block_mvp(torch.tensor(S, device="cuda")).cpu().numpy()
            ref = okr.block_mvp(n, 3, Ao, idx, diag, S, jpre)
            ops.append(f"block L={L}")
            assert rel(got, ref) < (1e-12 if address.size > 1 else 5e-12), ops
        elif op == "cuboid":
            L, prec = int(rng.integers(1, 41)), str(rng.choice(["fp64", "mixed"]))
            P = rng.standard_normal((3 * ncell, L)) + 1j * rng.standard_normal((3 * ncell, L))
            got = p.mvp_offdiag(torch.tensor(P, device="cuda"), prec).cpu().numpy()
            ops.append(f"cuboid L={L} {prec}")
            assert rel(got, omvp.fft_mvp(n, 3, Ao, P)) < (1e-12 if prec == "fp64" else 2e-6), ops
        elif op == "green":
            k = float(rng.uniform(5.0, 12.0))
            p.green_init(lf, k)
            Ao = omvp.fft_init(n, 3, lf, k)
            ops.append(f"green k={k:.3f}")
        else:
            lb = int(rng.choice([0, 1, 2, 3, 4, 8]))
            p.set_column_tile(lb)
            ops.append(f"tile {lb}")
    say(f"plan sequence n={n}: " + "; ".join(ops))


@pytest.mark.parametrize("n", [(1, 1, 1), (1, 1, 9), (1, 7, 1), (5, 1, 3), (2, 2, 2), (1, 30, 2)])
def test_one_cell_axes_match_oracle(n):
    """Cuboids with one-cell axes (a single dipole, a line, a plate): the
    doubled grid's length-2 axes and the zero planes of the Green tensor
    (exactly zero for one cell) on the occupied-block and cuboid products,
    both precisions, against the oracle."""
    say(f"one-cell axes n={n}")
    from bl_krylov import bl_krylov as pk
    from mvp_fft import mvp_fft as pm
    lf, k = 0.03, 9.0
    rng = np.random.default_rng(31)
    ncell = int(np.prod(n))
    L = 5
    P = rng.standard_normal((3 * ncell, L)) + 1j * rng.standard_normal((3 * ncell, L))
    Ao = omvp.fft_init(n, 3, lf, k)
    ref = omvp.fft_mvp(n, 3, Ao, P)
    At = pm.fft_init(n, 3, lf, k)
    got, gotm = pm.fft_mvp(n, 3, At, P), pm.fft_mvp(n, 3, At, P, precision="mixed")
    if ncell == 1:
        assert np.array_equal(ref, np.zeros_like(ref)) and np.array_equal(got, ref) and np.array_equal(gotm, ref)
    else:
        assert rel(got, ref) < 1e-12 and rel(gotm, ref) < 2e-6
    address = np.arange(ncell, dtype=np.int64)
    diag = rng.standard_normal(3 * ncell) + 1j * rng.standard_normal(3 * ncell) + 3.0
    jpre = 1.0 / diag
    S = rng.standard_normal((3 * ncell, L)) + 1j * rng.standard_normal((3 * ncell, L))
    idx = pk._build_scatter_index(address, 3)
    a = pk._block_mvp(n, 3, At, idx, diag, S, jpre)
    b = okr.block_mvp(n, 3, Ao, idx, diag, S, jpre)
    assert rel(a, b) < (1e-12 if ncell > 1 else 5e-12)


@pytest.mark.parametrize("n,L", [((36, 4, 30), 23), ((64, 5, 4), 17), ((64, 5, 4), 37), ((3, 4, 96), 29)])
def test_prime_block_width_runs_one_column_tiles(n, L):
    """A prime L above the tile cap leaves a 1-column tile (Q = 3), whose
    x pass stages more Green entries than line data; the launcher must keep
    it inside the CU's LDS (it once asked 183 KB on (36, 4, 30), L = 23).
    Against the oracle's _block_mvp, whole cuboid occupied."""
    say(f"prime block width n={n} L={L}")
    from bl_krylov import bl_krylov as pk
    from mvp_fft import mvp_fft as pm
    from bdda_amd.plan import Plan
    assert Plan(n).column_tile(L) == 1
    lf, k = 0.03, 9.5
    rng = np.random.default_rng(29)
    ncell = int(np.prod(n))
    address = np.arange(ncell, dtype=np.int64)
    diag = rng.standard_normal(3 * ncell) + 1j * rng.standard_normal(3 * ncell) + 3.0
    jpre = 1.0 / diag
    S = rng.standard_normal((3 * ncell, L)) + 1j * rng.standard_normal((3 * ncell, L))
    idx = pk._build_scatter_index(address, 3)
    a = pk._block_mvp(n, 3, pm.fft_init(n, 3, lf, k), idx, diag, S, jpre)
    b = okr.block_mvp(n, 3, omvp.fft_init(n, 3, lf, k), idx, diag, S, jpre)
    assert rel(a, b) < 1e-12


def test_block_mvp_after_other_products_on_the_same_plan():
    """The occupied operator keeps zeros on T1's unoccupied z lines and
    skips them; cuboid products (every line written), the complex64
    transform and another column count on the same plan in between must not
    leave it reading stale lines."""
    say("block mvp interleaved with cuboid products")
    from bl_krylov import bl_krylov as pk
    from mvp_fft import mvp_fft as pm
    n, box = (20, 18, 16), (9, 6, 16)
    lf, k = 0.03, 9.5
    rng = np.random.default_rng(23)
    cells = np.array([(ix * n[1] + iy) * n[2] + iz for ix in range(box[0]) for iy in range(box[1])
                      for iz in range(box[2])])
    address = np.sort(rng.choice(cells, size=cells.size // 4, replace=False)).astype(np.int64)
    N = address.size
    diag = rng.standard_normal(3 * N) + 1j * rng.standard_normal(3 * N) + 3.0
    jpre = 1.0 / diag
    idx = pk._build_scatter_index(address, 3)
    At = pm.fft_init(n, 3, lf, k)
    Ao = omvp.fft_init(n, 3, lf, k)
    ncell = int(np.prod(n))
    for L in (8, 5, 8):
        S = rng.standard_normal((3 * N, L)) + 1j * rng.standard_normal((3 * N, L))
        ref = okr.block_mvp(n, 3, Ao, idx, diag, S, jpre)
        assert rel(pk._block_mvp(n, 3, At, idx, diag, S, jpre), ref) < 1e-12
        P = rng.standard_normal((3 * ncell, L)) + 1j * rng.standard_normal((3 * ncell, L))
        pm.fft_mvp(n, 3, At, P)
        pm.fft_mvp(n, 3, At, P, precision="mixed")
        assert rel(pk._block_mvp(n, 3, At, idx, diag, S, jpre), ref) < 1e-12


def test_column_tile_prefers_the_compiled_8_column_tile():
    """Among equally padded tiles the operator takes 8 columns (the register
    kernels' compile-time tile); a tile that pads L less wins otherwise."""
    from bdda_amd.plan import Plan
    assert Plan((32, 9, 4)).column_tile(16) == 8   # 16 columns would also fit
    assert Plan((64, 6, 5)).column_tile(64) == 8
    assert Plan((64, 6, 5)).column_tile(70) == 10  # no padding
    assert Plan((64, 6, 5)).column_tile(5) == 5


def test_wide_column_tile_on_register_lengths():
    """L = 64 on a small grid selects 16-column tiles (Q = 48); lengths 48
    (16 x 3) and 80 (16 x 5) would need 768 threads in the register kernel
    and must take the Stockham kernel instead."""
    say("wide tiles")
    from mvp_fft import mvp_fft as pm
    n, L = (24, 3, 5), 64
    lf, k = 0.03, 9.0
    rng = np.random.default_rng(5)
    ncell = int(np.prod(n))
    P = rng.standard_normal((3 * ncell, L)) + 1j * rng.standard_normal((3 * ncell, L))
    ref = omvp.fft_mvp(n, 3, omvp.fft_init(n, 3, lf, k), P)
    At = pm.fft_init(n, 3, lf, k)
    assert rel(pm.fft_mvp(n, 3, At, P), ref) < 1e-12
    assert rel(pm.fft_mvp(n, 3, At, P, precision="mixed"), ref) < 2e-6


def test_green_spectrum_and_tensor_match_oracle():
    say("green spectrum")
    from mvp_fft import mvp_fft as pm
    n, lf, k = (5, 6, 4), 0.04, 9.1
    At = pm.fft_init(n, 3, lf, k)
    assert At.shape == (10, 12, 8, 3, 3)
    g = np.asarray(At)
    ref = omvp.fft_init(n, 3, lf, k)
    assert rel(g, ref) < 1e-13
    t = pm.build_interaction_tensor(n, lf, k)
    tref = omvp.interaction_tensor(n, lf, k)
    assert rel(t, tref) < 1e-14
    assert np.array_equal(t == 0, tref == 0)  # identical zero planes
    with pytest.raises(ValueError):
        pm.fft_init(n, 2, lf, k)


def test_green_non_smooth_grid_still_exact_convolution():
    say("non-smooth grid")
    from mvp_fft import mvp_fft as pm
    n = (17, 7, 13)  # 2N = 34, 14, 26 are not 5-smooth
    At = pm.fft_init(n, 3, 0.02, 15.0)
    assert At.shape[:3] == (36, 15, 25)
    with pytest.raises(ValueError):
        np.asarray(At)


# ----------------------------------------------------------------- solver
def _problem(r=0.07, bc=1.2, ab=1.1, beta=0.0, m_p=(1.5 + 0.01j,) * 3, L=6, wl=0.35, mm=1.2, seed=0):
    name, n, lf, grid, is_in, rve, _ = ogre.build_gre_geometry(r, bc, ab, beta, seed=seed)
    return name, n, lf, grid, is_in, np.asarray(m_p), orient.euler_run_dda(L, seed=seed + 1), wl, mm


def _oracle_solve(prob, solver, tol, itermax):
    name, n, lf, grid, is_in, m_p, eu, wl, mm = prob
    tg = osc.Target(n, lf, grid, is_in, m_p)
    inc = osc.Incident(wl, mm, eu)
    return osc.solve(tg, inc, solver, tol, itermax)


def _gpu_dd(prob, solver="bicgstab", tol=1e-2, itermax=25, precision="fp64", **kw):
    from bl_dda.scatterer import DiscreteDipoles, IncidentField, Target
    name, n, lf, grid, is_in, m_p, eu, wl, mm = prob
    dd = DiscreteDipoles(Target(name, n, lf, grid, is_in, m_p), IncidentField(wl, mm, eu))
    dd.verbose = False
    dd.solver, dd.tol, dd.itermax, dd.precision = solver, tol, itermax, precision
    for a, v in kw.items():
        setattr(dd, a, v)
    dd.set_interaction_matrix()
    dd.solve_matrix_equation()
    return dd


@pytest.mark.parametrize("solver", ["cocg_rq", "bicgstab"])
@pytest.mark.parametrize("L", [150, 200])
def test_wide_blocks_match_oracle(solver, L):
    """More orientations than 128 in one block solve (the reference has no
    limit): three and four 64-column Gram tiles with fewer row parts, the
    generic update kernel.  As for L = 70, ill-conditioned L x L systems
    let rounding differences grow, so the converged solutions agree to the
    tolerance rather than iterate for iterate."""
    say(f"wide block {solver} L={L}")
    prob = _problem(r=0.12, L=L)
    ref = _oracle_solve(prob, "cocg" if solver == "cocg_rq" else "bicgstab", 1e-8, 80)
    dd = _gpu_dd(prob, solver, 1e-8, 80)
    assert dd.converge and abs(dd.iter_fin - ref["iter_fin"]) <= 2
    assert rel(dd.X, ref["X"]) < 1e-6 and dd.true_residual < 1e-7
    np.testing.assert_allclose(dd.compute_C_ext(), ref["C_ext"], rtol=1e-6)
    sb = dd.compute_OCBS_observable_S_bk()
    assert np.max(np.abs(sb - ref["S_bk"]) / np.abs(ref["S_bk"])) < 1e-6


@pytest.mark.parametrize("draw", range(10))
def test_random_particles_solve_as_the_oracle(draw):
    """Seeded random solves: GRE particles of 300-700 dipoles (size, ratios,
    beta, seed drawn), isotropic or birefringent m_p, L of 1-40 (prime
    widths above the tile cap take 1-column tiles), either solver (block-
    COCG-rq only on isotropic m_p, whose preconditioned operator is complex
    symmetric, its condition in the reference), the host or the device
    L x L algebra, float64 or mixed.  Where the oracle's float64
    solve converges to 1e-8 the GPU solve must as well, with the same
    solution to 1e-6 and S, C within 1e-6 (float64), or S within 1e-4 at a
    true residual of 1e-5 (mixed).  Where the oracle does not converge,
    only the GPU's own report is checked."""
    from bdda_amd import _lib
    rng = np.random.default_rng(2000 + draw)
    r = float(rng.uniform(0.05, 0.10))
    bc, ab = float(rng.uniform(0.6, 1.6)), float(rng.uniform(0.6, 1.4))
    beta, seed = float(rng.choice([0.0, 0.1])), int(rng.integers(0, 1000))
    iso = bool(rng.integers(0, 2))
    if iso:
        m_p = (complex(rng.uniform(1.3, 1.7), 10 ** rng.uniform(-4, -1.5)),) * 3
    else:
        m_p = tuple(complex(rng.uniform(1.3, 1.7), 10 ** rng.uniform(-4, -1.5)) for _ in range(3))
    L, mm = int(rng.integers(1, 41)), float(rng.uniform(1.0, 1.33))
    # block-COCG-rq needs the complex-symmetric operator of an isotropic m_p
    solver = str(rng.choice(["cocg_rq", "bicgstab"])) if iso else "bicgstab"
    la, precision = int(rng.integers(0, 2)), str(rng.choice(["fp64", "mixed"]))
    say(f"random solve draw={draw} L={L} {solver} la={la} {precision}")
    prob = _problem(r=r, bc=bc, ab=ab, beta=beta, m_p=m_p, L=L, mm=mm, seed=seed)
    ref = _oracle_solve(prob, "cocg" if solver == "cocg_rq" else "bicgstab", 1e-8, 300)
    lib = _lib.load()
    try:
        _lib.check(lib.bdda_set_la_path(la))
        dd = _gpu_dd(prob, solver, 1e-8 if precision == "fp64" else 1e-5, 300, precision)
    finally:
        lib.bdda_set_la_path(0)
    assert dd.converge == (dd.err_fin < dd.tol)
    if not ref["err_fin"] < 1e-8:
        return
    assert dd.converge
    st, sp = dd.compute_PCAS_observable_S_fw()
    sb = dd.compute_OCBS_observable_S_bk()
    if precision == "fp64":
        assert rel(dd.X, ref["X"]) < 1e-6 and dd.true_residual < 1e-7
        for a, b in [(st, ref["S_fw_theta"]), (sp, ref["S_fw_phi"]), (sb, ref["S_bk"])]:
            assert np.max(np.abs(a - b) / np.abs(b)) < 1e-6
        np.testing.assert_allclose(dd.compute_C_ext(), ref["C_ext"], rtol=1e-6)
    else:
        assert dd.true_residual <= 1e-5
        for a, b in [(st, ref["S_fw_theta"]), (sp, ref["S_fw_phi"]), (sb, ref["S_bk"])]:
            assert np.max(np.abs(a - b) / np.abs(b)) < 1e-4


def test_block_width_limit_is_reported():
    from bdda_amd.plan import Plan
    p = Plan((4, 4, 4))
    p.green_init(0.03, 9.0)
    addr = np.arange(64)
    p.set_target(addr, np.full(192, 3.0 + 0j), np.full(192, 1 / 3.0 + 0j))
    B = torch.zeros((192, 4097), dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError, match="4096"):
        p.solve(B)


@pytest.mark.parametrize("solver", ["cocg_rq", "bicgstab"])
@pytest.mark.parametrize("L", [6, 24, 37, 64, 70])
def test_solver_iterates_and_observables_match_oracle(solver, L):
    """Block solves against the oracle; L = 6, 24 and 70 put the Gram and
    update kernels on their 16-, 32- and 64-wide tiles (two of them at 70);
    37 and 64 take the matrix-core Gram and update on one 64-wide tile (37:
    partial coefficient steps and columns)."""
    say(f"solver {solver} L={L}")
    prob = _problem(L=L)
    ref = _oracle_solve(prob, "cocg" if solver == "cocg_rq" else "bicgstab", 1e-8, 60)
    dd = _gpu_dd(prob, solver, 1e-8, 60)
    if L <= 8:
        # same iterate sequence as the oracle
        assert dd.iter_fin == ref["iter_fin"]
        np.testing.assert_allclose(dd.err_history, ref["history"], rtol=1e-6)
        assert rel(dd.X, ref["X"]) < 1e-9
    else:
        # 24 or 70 orientations on a ~500-dipole target make the L x L block
        # coefficient systems ill-conditioned, so rounding differences between
        # any two implementations grow along the iteration (the reference has
        # the same property); the converged solutions agree to the tolerance
        assert abs(dd.iter_fin - ref["iter_fin"]) <= 1
        np.testing.assert_allclose(dd.err_history[0], ref["history"][0], rtol=1e-4)
        assert rel(dd.X, ref["X"]) < 1e-6
    st, sp = dd.compute_PCAS_observable_S_fw()
    sb = dd.compute_OCBS_observable_S_bk()
    for a, b in [(st, ref["S_fw_theta"]), (sp, ref["S_fw_phi"]), (sb, ref["S_bk"])]:
        assert np.max(np.abs(a - b) / np.abs(b)) < 1e-6
    np.testing.assert_allclose(dd.compute_C_ext(), ref["C_ext"], rtol=1e-6)
    np.testing.assert_allclose(dd.compute_C_abs(), ref["C_abs"], rtol=1e-6)
    assert dd.true_residual < 1e-7


@pytest.mark.parametrize("solver", ["cocg_rq", "bicgstab"])
@pytest.mark.parametrize("L", [6, 24, 64])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_and_host_la_paths_agree(solver, L, precision):
    """The L x L recurrence algebra on the host (default) and on the device
    (bdda_set_la_path(1), L <= 64) run the same recurrences: same
    iteration count at L = 6, the first residuals equal to rounding, the
    solutions equal to the tolerance class (rounding grows along
    ill-conditioned iterations at L = 24 and 64)."""
    say(f"la paths {solver} L={L} {precision}")
    from bdda_amd import _lib
    lib = _lib.load()
    prob = _problem(L=L)
    out = {}
    try:
        for path in (0, 1):
            _lib.check(lib.bdda_set_la_path(path))
            out[path] = _gpu_dd(prob, solver, 1e-8, 80, precision=precision)
    finally:
        lib.bdda_set_la_path(0)
    h, d = out[0], out[1]
    assert h.converge and d.converge and d.true_residual < 1e-7
    # the first iteration's L x L systems are already ill-conditioned at
    # L = 24, 64 on ~500 dipoles: rounding differences reach ~1e-10 there
    np.testing.assert_allclose(d.err_history[0], h.err_history[0], rtol=1e-10 if L <= 8 else 1e-7)
    if L <= 8:
        assert d.iter_fin == h.iter_fin
        # in mixed mode the residuals after the complex64 cycle carry that
        # cycle's product rounding (about 1e-7 relative per product): the
        # paths then agree to 1e-12 in the relative residual, not to 1e-6 of it
        np.testing.assert_allclose(d.err_history, h.err_history, rtol=1e-6,
                                   atol=1e-12 if precision == "mixed" else 0.0)
        assert rel(d.X, h.X) < 1e-9
    else:
        assert abs(d.iter_fin - h.iter_fin) <= 1
        assert rel(d.X, h.X) < 1e-6


def test_la_path_rejects_unknown():
    from bdda_amd import _lib
    assert _lib.load().bdda_set_la_path(2) == _lib.E_ARG


def test_block_mvp_and_krylov_dropins_match_oracle():
    say("block_mvp drop-in")
    from bl_krylov import bl_krylov as pk
    from mvp_fft import mvp_fft as pm
    name, n, lf, grid, is_in, m_p, eu, wl, mm = _problem(L=4)
    tg = osc.Target(n, lf, grid, is_in, m_p)
    inc = osc.Incident(wl, mm, eu)
    eps_r, alpha, diag = osc.polarizability(tg, inc)
    E_inc, B = osc.rhs(tg, inc)
    jpre = 1.0 / diag
    At_o = omvp.fft_init(n, 3, lf, inc.k)
    At_p = pm.fft_init(n, 3, lf, inc.k)
    idx = pk._build_scatter_index(tg.address, 3)
    rng = np.random.default_rng(3)
    S = rng.standard_normal(B.shape) + 1j * rng.standard_normal(B.shape)
    a = pk._block_mvp(n, 3, At_p, idx, diag, S, jpre)
    b = okr.block_mvp(n, 3, At_o, idx, diag, S, jpre)
    assert rel(a, b) < 1e-12
    X, it, err = pk.bl_bicgstab_jacobi_mvp_fft(n, 3, tg.address, At_p, diag, B, 1e-6, 30, verbose=False)
    Xo, ito, erro, _ = okr.bl_bicgstab(n, 3, tg.address, At_o, diag, B, 1e-6, 30)
    assert it == ito and abs(err - erro) < 1e-6 * erro and rel(X, Xo) < 1e-9
    X, it, err = pk.bl_cocg_rq_jacobi_mvp_fft(n, 3, tg.address, At_p, diag, B, 1e-6, 30, verbose=False)
    Xo, ito, erro, _ = okr.bl_cocg_rq(n, 3, tg.address, At_o, diag, B, 1e-6, 30)
    assert it == ito and abs(err - erro) < 1e-6 * erro and rel(X, Xo) < 1e-9


def test_notebook_run_on_gpu(golden_nb):
    """test_dda.ipynb: sphere r_v = 0.2 um, 17^3 lattice, 10 orientations."""
    say("notebook")
    from bl_dda.scatterer import DiscreteDipoles, IncidentField, Target
    from shape_model.gaussian_ellipsoid import gaussian_ellipsoid_shape_model as G
    nb = golden_nb
    g = G(float(nb["r_v_base"]), 1, 1, 0.0)
    pts, _ = g.compute_r_points_on_GRE(np.random.default_rng(int(nb["shape_seed"])))
    _, n, grid = g.create_cuboid_lattice_that_encloses_GRE_shape(pts)
    dist = g.find_nearest_distance_from_the_GRE_surf(grid, pts)
    is_in = g.extract_lattice_address_in_GRE_volume(g.lattice_lf, g.distance_factor, n, dist)
    assert n.tolist() == [17, 17, 17]
    tg = Target(g.name, n, g.lattice_lf, grid, is_in, nb["m_p_xyz"])
    dd = DiscreteDipoles(tg, IncidentField(float(nb["wl_0"]), float(nb["m_m"]), orient.euler_test_dda()))
    assert dd.dpl == float(nb["dpl"])
    dd.verbose = False
    dd.set_interaction_matrix()
    dd.solve_matrix_equation()
    assert dd.converge and dd.iter_fin == int(nb["iter_fin"])
    assert ["%.4f" % e for e in dd.err_history] == ["%.4f" % e for e in nb["err_printed"]]
    np.testing.assert_allclose(dd.compute_OCBS_observable_S_bk(), nb["s_bk"], rtol=2e-6, atol=1e-8)
    # cells 9-12: the volume-equivalent sphere's Mie values through the
    # drop-in Mie module; the recorded per-orientation errors of Q_abs, Q_ext
    # against Mie, and the notebook's 1% tolerance on S(0)
    from analytical_scattering_theories.homogeneous_sphere import mie_compute_q_and_s
    m_p_avg = (1.5 + 0.01j) * 3 / 3
    q = mie_compute_q_and_s(float(nb["wl_0"]), float(nb["m_m"]), dd.ve_radius, m_p_avg, nang=3)
    np.testing.assert_allclose(q[:3], nb["mie_q"], rtol=1e-13)
    qabs = dd.compute_C_abs() / (np.pi * dd.ve_radius ** 2)
    qext = dd.compute_C_ext() / (np.pi * dd.ve_radius ** 2)
    np.testing.assert_allclose(np.abs(qabs / q[1] - 1), nb["qabs_relerr"], rtol=5e-5, atol=1e-8)
    np.testing.assert_allclose(np.abs(qext / q[2] - 1), nb["qext_relerr"], rtol=5e-5, atol=1e-8)
    st, sp = dd.compute_PCAS_observable_S_fw()
    for s in (st, sp):
        assert np.all(np.abs(s.real / q[3].real - 1) < 0.01) and np.all(np.abs(s.imag / q[3].imag - 1) < 0.01)


def test_mixed_precision_meets_fp64_oracle():
    say("mixed precision")
    # BiCGSTAB on a birefringent particle; COCG-rq needs a complex-symmetric
    # preconditioned operator, i.e. isotropic m_p (the reference's condition)
    cases = [("bicgstab", (1.55 + 0.001j, 1.6 + 0.001j, 1.5 + 0.001j)), ("cocg_rq", (1.5 + 0.01j,) * 3)]
    for solver, m_p in cases:
        prob = _problem(r=0.09, L=8, m_p=m_p)
        ref = _oracle_solve(prob, "bicgstab", 1e-10, 80)
        dd = _gpu_dd(prob, solver, 1e-5, 80, precision="mixed", refine_every=20)
        assert dd.converge and dd.true_residual < 1e-5
        st, sp = dd.compute_PCAS_observable_S_fw()
        sb = dd.compute_OCBS_observable_S_bk()
        for a, b in [(st, ref["S_fw_theta"]), (sp, ref["S_fw_phi"]), (sb, ref["S_bk"])]:
            assert np.max(np.abs(a - b) / np.abs(b)) < 1e-4


def test_warm_start_and_fallback():
    say("warm start")
    prob = _problem(L=5)
    ref = _oracle_solve(prob, "bicgstab", 1e-9, 60)
    good = ref["X"] * (1 + 1e-4)
    dd = _gpu_dd(prob, "bicgstab", 1e-5, 60, X0=good)
    assert dd.warm_start_used and dd.converge and dd.iter_fin <= 1
    bad = ref["X"] * 50.0
    dd2 = _gpu_dd(prob, "bicgstab", 1e-5, 60, X0=bad)
    dd0 = _gpu_dd(prob, "bicgstab", 1e-5, 60)
    assert not dd2.warm_start_used and dd2.iter_fin == dd0.iter_fin
    assert rel(dd2.X, dd0.X) < 1e-12


@pytest.mark.parametrize("draw", range(4))
def test_warm_start_on_random_guesses(draw):
    """Seeded random guesses X0 = X (1 + e Z) around the oracle's solution,
    e from 1e-6 to 30, on a random particle, width, solver and precision.
    A guess whose preconditioned residual is below the zero start's is taken
    and the solve reaches the tolerance from it (in fewer iterations for a
    close guess); a worse one is dropped, and the solve then takes the
    zero start's iterations to the same solution."""
    rng = np.random.default_rng(7000 + draw)
    L = int(rng.integers(1, 30))
    solver, precision = str(rng.choice(["cocg_rq", "bicgstab"])), str(rng.choice(["fp64", "mixed"]))
    prob = _problem(r=float(rng.uniform(0.05, 0.09)), L=L, seed=int(rng.integers(0, 1000)))
    ref = _oracle_solve(prob, "cocg" if solver == "cocg_rq" else "bicgstab", 1e-10, 200)
    dd0 = _gpu_dd(prob, solver, 1e-5, 200, precision)
    for e in (1e-6, 1e-3, 0.1, 3.0, 30.0):
        Z = rng.standard_normal(ref["X"].shape) + 1j * rng.standard_normal(ref["X"].shape)
        X0 = ref["X"] * (1 + e * Z)
        say(f"warm start draw={draw} L={L} {solver} {precision} e={e}")
        dd = _gpu_dd(prob, solver, 1e-5, 200, precision, X0=X0)
        assert dd.converge and dd.true_residual <= 1e-5
        if dd.warm_start_used:
            if e <= 1e-3:
                assert dd.iter_fin < dd0.iter_fin
        else:
            assert dd.iter_fin == dd0.iter_fin and rel(dd.X, dd0.X) < 1e-12


@pytest.mark.parametrize("draw", range(8))
def test_solver_limits_on_random_draws(draw):
    """Seeded random tolerances (1e-1 to 1e-9, above the rounding floor) and
    iteration budgets (0 to 40) at small L, where the float64 iterates
    follow the oracle's: the same iteration count, residual history,
    reported error and convergence flag, including a budget of 0 (no
    iteration, error inf, not converged) and a budget that runs out.  The
    mixed solve under the same limits (refinement checks every 1-20
    iterations or never) must report a consistent state: converged exactly
    when its error, its true residual, is below the tolerance."""
    rng = np.random.default_rng(8000 + draw)
    L = int(rng.integers(1, 9))
    solver = str(rng.choice(["cocg_rq", "bicgstab"])) if L > 1 else "bicgstab"
    tol = float(rng.choice([1e-1, 1e-3, 1e-6, 1e-9]))
    itermax = int(rng.choice([0, 1, 2, 5, 40]))
    every = int(rng.choice([0, 1, 3, 20]))
    say(f"solver limits draw={draw} L={L} {solver} tol={tol} itermax={itermax} refine_every={every}")
    prob = _problem(r=float(rng.uniform(0.05, 0.08)), L=L, seed=int(rng.integers(0, 1000)))
    ref = _oracle_solve(prob, "cocg" if solver == "cocg_rq" else "bicgstab", tol, itermax)
    dd = _gpu_dd(prob, solver, tol, itermax)
    assert dd.iter_fin == ref["iter_fin"] and dd.converge == ref["converge"]
    if itermax == 0:
        assert dd.err_fin == np.inf and ref["err_fin"] == np.inf and not dd.converge
    else:
        np.testing.assert_allclose(dd.err_history, ref["history"], rtol=1e-6)
        np.testing.assert_allclose(dd.err_fin, ref["err_fin"], rtol=1e-6)
    dm = _gpu_dd(prob, solver, tol, itermax, "mixed", refine_every=every)
    assert dm.converge == (dm.err_fin < tol)
    if dm.converge:
        assert dm.true_residual < tol
    assert len(dm.err_history) <= max(itermax, 1)


@pytest.mark.parametrize("draw", range(3))
def test_guess_projection_is_the_least_squares_fit(draw):
    """bdda_project_guess replaces X0 by X0 C with C the least-squares fit
    min ||jpre (B - A X0 C)||_F (normal equations with a relative ridge of
    1e-13 tr(W^H W) / L, W = jpre A X0).  Against the same fit in NumPy on
    the oracle's operator: a noisy guess, the exact solution with its
    columns mixed (the fit must undo the mixing), and a guess with two
    equal columns (the ridge keeps the normal equations solvable)."""
    rng = np.random.default_rng(9000 + draw)
    L = int(rng.integers(3, 20))
    prob = _problem(r=float(rng.uniform(0.05, 0.08)), L=L, seed=int(rng.integers(0, 1000)))
    ref = _oracle_solve(prob, "bicgstab", 1e-11, 200)
    dd = _gpu_dd(prob, "bicgstab", 1e-2, 1)
    name, n, lf, grid, is_in, m_p, eu, wl, mm = prob
    address = np.flatnonzero(is_in.ravel()) if is_in.ndim > 1 else np.flatnonzero(is_in)
    idx = okr.scatter_index(address, 3)
    diag, B, At = ref["diag"], ref["B"], ref["A_tilde"]
    jpre = 1.0 / diag
    Bj = B * jpre[:, None]
    X = ref["X"]
    Z = rng.standard_normal(X.shape) + 1j * rng.standard_normal(X.shape)
    M = rng.standard_normal((L, L)) + 1j * rng.standard_normal((L, L)) + 3 * np.eye(L)
    dup = X * (1 + 0.1 * Z)
    dup[:, 1] = dup[:, 0]
    for kind, X0 in (("noisy", X * (1 + 0.3 * Z)), ("mixed columns", X @ M), ("equal columns", dup)):
        say(f"guess projection draw={draw} L={L} {kind}")
        W = okr.block_mvp(n, 3, At, idx, diag, X0, jpre)
        G = W.conj().T @ W
        G = G + 1e-13 * np.trace(G).real / L * np.eye(L)
        C = np.linalg.solve(G, W.conj().T @ Bj)
        want_before = np.linalg.norm(Bj - W) / np.linalg.norm(Bj)
        want_after = np.linalg.norm(Bj - W @ C) / np.linalg.norm(Bj)
        X0d = torch.tensor(X0, dtype=torch.complex128, device="cuda").contiguous()
        before, after = dd._plan.project_guess(dd._B_dev, X0d)
        got = X0d.cpu().numpy()
        assert np.isfinite(got).all() and after <= before
        np.testing.assert_allclose(before, want_before, rtol=1e-9)
        np.testing.assert_allclose(after, want_after, rtol=1e-6, atol=1e-10)
        assert rel(got, X0 @ C) < 1e-8
        if kind == "mixed columns":
            # up to the ridge's bias, ~1e-13 cond(W)^2
            assert after < 1e-7 and rel(got, X) < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_warm_start_that_already_converged(precision):
    """A guess that already meets the tolerance: the float64 solve runs the
    reference's one iteration; the mixed solve needs no complex64 cycle and
    reports the guess's true residual as its one history entry (it once
    reported no history, so the solve read as not converged)."""
    say(f"converged warm start {precision}")
    prob = _problem(L=6)
    ref = _oracle_solve(prob, "bicgstab", 1e-10, 80)
    dd = _gpu_dd(prob, "bicgstab", 1e-5, 80, precision, X0=ref["X"])
    assert dd.warm_start_used and dd.converge and dd.iter_fin == 0
    assert len(dd.err_history) == 1 and dd.err_fin < 1e-5 and dd.true_residual < 1e-5
    assert rel(dd.X, ref["X"]) < 1e-6


def test_cocg_nearly_dependent_orientations():
    """Two orientations 1e-9 rad apart make the right-hand-side block nearly
    rank deficient (its Gram matrix is singular to float64, so a plain
    Cholesky QR fails); the shifted QR keeps block-COCG-rq going and every
    column still matches its own single-orientation solve.  The single
    solves use BiCGSTAB: with one circularly polarised right-hand side
    b^T b = 0 (e0^T e0 = 0), so COCG at L = 1 breaks down (Q^T Q = 0) in the
    reference as in this solver."""
    say("nearly dependent orientations")
    prob = list(_problem(L=4))
    eu = np.array(prob[6], dtype=np.float64)
    eu[1] = eu[0] + 1e-9
    prob[6] = eu
    from bdda_amd import _lib
    lib = _lib.load()
    try:  # the shifted QR on the device's and on the host's L x L algebra
        _lib.check(lib.bdda_set_la_path(1))
        dh = _gpu_dd(tuple(prob), "cocg_rq", 1e-9, 80)
    finally:
        lib.bdda_set_la_path(0)
    assert dh.converge and dh.true_residual < 1e-8
    dd = _gpu_dd(tuple(prob), "cocg_rq", 1e-9, 80)
    assert dd.converge and dd.true_residual < 1e-8
    assert rel(dd.X, dh.X) < 1e-7
    for l in (0, 1, 3):
        one = list(prob)
        one[6] = eu[l:l + 1]
        d1 = _gpu_dd(tuple(one), "bicgstab", 1e-10, 80)
        assert rel(dd.X[:, l], d1.X[:, 0]) < 1e-7


@pytest.mark.parametrize("la_path", [0, 1])
def test_repeated_orientations_as_the_reference(la_path):
    """Four orientations given twice make the right-hand-side block exactly
    rank deficient.  The reference's block-COCG-rq proceeds (numpy's
    Householder QR gives a zero row in R) and solves both copies alike; its
    block-BiCGSTAB raises LinAlgError (singular R0^H V).  Here even shifted
    Cholesky QR breaks down on such a block, and the Householder QR on the
    host takes over, on either L x L algebra path."""
    say(f"repeated orientations la_path={la_path}")
    prob = list(_problem(L=16))
    eu = np.concatenate([prob[6], prob[6][:4]])
    prob[6] = eu
    prob = tuple(prob)
    ref = _oracle_solve(prob, "cocg", 1e-8, 80)
    from bdda_amd import _lib
    lib = _lib.load()
    try:
        _lib.check(lib.bdda_set_la_path(la_path))
        dd = _gpu_dd(prob, "cocg_rq", 1e-8, 80)
        with pytest.raises(np.linalg.LinAlgError):
            _gpu_dd(prob, "bicgstab", 1e-8, 80)
    finally:
        lib.bdda_set_la_path(0)
    assert dd.converge and dd.true_residual < 1e-7
    assert abs(dd.iter_fin - ref["iter_fin"]) <= 1 and rel(dd.X, ref["X"]) < 1e-6
    assert rel(dd.X[:, 16:], dd.X[:, :4]) < 1e-8
    np.testing.assert_allclose(dd.compute_C_ext(), ref["C_ext"], rtol=1e-6)


def test_nonconvergence_is_reported_not_raised():
    say("non-convergence")
    prob = _problem(L=3)
    dd = _gpu_dd(prob, "bicgstab", 1e-30, 2)
    assert not dd.converge and dd.iter_fin == 1 and dd.err_fin > 1e-30
    assert np.all(dd.P == 0)
    dd0 = _gpu_dd(prob, "cocg_rq", 1e-2, 0)
    assert not dd0.converge and dd0.err_fin == float("inf")


def test_single_dipole():
    say("single dipole")
    from bl_dda.scatterer import DiscreteDipoles, IncidentField, Target
    grid = np.zeros((1, 3))
    tg = Target("one", np.array([1, 1, 1]), 0.02, grid, np.array([True]), np.array([1.6 + 0.1j] * 3))
    dd = DiscreteDipoles(tg, IncidentField(0.5, 1.1, orient.euler_run_dda(3, seed=4)))
    dd.verbose = False
    dd.tol = 1e-12
    dd.set_interaction_matrix()
    dd.solve_matrix_equation()
    a = dd.alpha_E[0, 0]
    np.testing.assert_allclose(dd.compute_PCAS_observable_S_fw()[0], dd.k ** 2 * a * np.ones(3), rtol=1e-13)


def test_padded_cuboid_is_the_same_physics():
    say("padding")
    prob = _problem(L=4)
    a = _gpu_dd(prob, "bicgstab", 1e-8, 60)
    b = _gpu_dd(prob, "bicgstab", 1e-8, 60, pad_to=(16, 16, 16))
    assert tuple(b.Au_til.n) == (16, 16, 16)
    assert a.iter_fin == b.iter_fin and rel(b.X, a.X) < 1e-10


@pytest.mark.parametrize("beta,seed", [(0.2, 3), (0.3, 8)])
def test_trimmed_cuboid_is_the_same_physics(beta, seed):
    """Without pad_to the solver's cuboid is the occupied extent of the
    lattice (here several cells short of it on every axis, so the origin
    moves too); the solve equals the one on the whole lattice to rounding
    and matches the oracle's, which runs on the whole lattice."""
    say(f"trimmed cuboid beta={beta} seed={seed}")
    prob = _problem(r=0.08, bc=0.8, ab=0.7, beta=beta, L=5, seed=seed)
    name, n, lf, grid, is_in, *_ = prob
    occ = np.argwhere(np.asarray(is_in).reshape(tuple(int(v) for v in n)))
    ext = tuple(int(v) for v in occ.max(0) - occ.min(0) + 1)
    assert ext != tuple(int(v) for v in n) and occ.min(0).max() > 0
    a = _gpu_dd(prob, "bicgstab", 1e-8, 80)
    b = _gpu_dd(prob, "bicgstab", 1e-8, 80, pad_to=tuple(int(v) for v in n))
    assert a.solver_dims == ext and b.solver_dims == tuple(int(v) for v in n)
    assert a.converge and a.iter_fin == b.iter_fin and rel(a.X, b.X) < 1e-10
    ref = _oracle_solve(prob, "bicgstab", 1e-8, 80)
    assert ref["iter_fin"] == a.iter_fin and rel(a.X, ref["X"]) < 1e-9
    np.testing.assert_allclose(a.compute_C_ext(), ref["C_ext"], rtol=1e-6)
    np.testing.assert_allclose(a.compute_OCBS_observable_S_bk(), ref["S_bk"], rtol=1e-6)


# -------------------------------------------------------------- voxeliser
@pytest.mark.parametrize("beta", [0.0, 0.2])
def test_voxeliser_bit_exact(beta):
    say(f"voxeliser beta={beta}")
    from shape_model.gaussian_ellipsoid import gaussian_ellipsoid_shape_model as G
    g = G(0.15, 1.3, 1.2, beta)
    pts, _ = g.compute_r_points_on_GRE(np.random.default_rng(11))
    _, n, grid = g.create_cuboid_lattice_that_encloses_GRE_shape(pts)
    d_gpu = g.find_nearest_distance_from_the_GRE_surf(grid, pts)
    d_ref = ogre.nearest_distance(grid, pts)
    assert np.array_equal(d_gpu, d_ref)
    m_gpu = g.extract_lattice_address_in_GRE_volume(g.lattice_lf, g.distance_factor, n, d_gpu)
    m_ref = ogre.inside_mask(g.lattice_lf, g.distance_factor, n, d_ref)
    assert np.array_equal(m_gpu, m_ref)


@pytest.mark.parametrize("draw", range(6))
def test_voxeliser_bit_exact_on_random_particles(draw):
    """Seeded random GRE particles (size, both ratios, beta up to 0.3, shape
    seed): the GPU nearest distances and inside mask equal the oracle's
    bit for bit."""
    from shape_model.gaussian_ellipsoid import gaussian_ellipsoid_shape_model as G
    rng = np.random.default_rng(5000 + draw)
    r, bc, ab = float(rng.uniform(0.06, 0.2)), float(rng.uniform(0.5, 1.6)), float(rng.uniform(0.5, 1.4))
    beta, seed = float(rng.uniform(0.0, 0.3)), int(rng.integers(0, 10000))
    say(f"voxeliser draw={draw} r={r:.3f} bc={bc:.2f} ab={ab:.2f} beta={beta:.2f} seed={seed}")
    g = G(r, bc, ab, beta)
    pts, _ = g.compute_r_points_on_GRE(np.random.default_rng(seed))
    _, n, grid = g.create_cuboid_lattice_that_encloses_GRE_shape(pts)
    d_gpu = g.find_nearest_distance_from_the_GRE_surf(grid, pts)
    d_ref = ogre.nearest_distance(grid, pts)
    assert np.array_equal(d_gpu, d_ref)
    m_gpu = g.extract_lattice_address_in_GRE_volume(g.lattice_lf, g.distance_factor, n, d_gpu)
    m_ref = ogre.inside_mask(g.lattice_lf, g.distance_factor, n, d_ref)
    assert np.array_equal(m_gpu, m_ref)


@pytest.mark.parametrize("groups", [2, 3, 4])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_operator_groups_are_bit_identical(groups, precision):
    """The occupied operator with its column tiles split into 2-4 groups on
    staggered streams computes every column as the plain operator does:
    block products and whole solves bit-identical (L = 40: five 8-column
    tiles, so groups of uneven size too)."""
    say(f"operator groups {groups} {precision}")
    from bdda_amd import _lib
    lib = _lib.load()
    prob = _problem(L=40)
    ref = _gpu_dd(prob, "cocg_rq", 1e-8, 80, precision=precision)
    g = torch.Generator(device="cuda").manual_seed(5)
    S = torch.randn(ref._B_dev.shape, dtype=torch.complex128, device="cuda", generator=g)
    AS_ref = ref._plan.block_mvp(S)
    try:
        _lib.check(lib.bdda_set_operator_groups(groups))
        AS = ref._plan.block_mvp(S)
        dd = _gpu_dd(prob, "cocg_rq", 1e-8, 80, precision=precision)
    finally:
        lib.bdda_set_operator_groups(1)
    assert torch.equal(AS, AS_ref)
    assert dd.iter_fin == ref.iter_fin and dd.err_history == ref.err_history
    assert torch.equal(dd.X_device, ref.X_device)


@pytest.mark.parametrize("draw", range(4))
def test_operator_groups_bit_identical_on_random_widths(draw):
    """Seeded random widths L of 1-90 (compiled and runtime tiles, from one
    tile to dozens, groups of uneven size or fewer tiles than groups), 2-4
    groups, either precision: the grouped occupied operator equals the
    plain one bit for bit, on the block product and on a whole solve."""
    from bdda_amd import _lib
    rng = np.random.default_rng(6000 + draw)
    L, groups = int(rng.integers(1, 91)), int(rng.integers(2, 5))
    precision = str(rng.choice(["fp64", "mixed"]))
    say(f"operator groups draw={draw} L={L} groups={groups} {precision}")
    lib = _lib.load()
    prob = _problem(r=float(rng.uniform(0.05, 0.09)), L=L, seed=int(rng.integers(0, 1000)))
    ref = _gpu_dd(prob, "bicgstab", 1e-8, 80, precision=precision)
    g = torch.Generator(device="cuda").manual_seed(draw)
    S = torch.randn(ref._B_dev.shape, dtype=torch.complex128, device="cuda", generator=g)
    AS_ref = ref._plan.block_mvp(S)
    try:
        _lib.check(lib.bdda_set_operator_groups(groups))
        AS = ref._plan.block_mvp(S)
        dd = _gpu_dd(prob, "bicgstab", 1e-8, 80, precision=precision)
    finally:
        lib.bdda_set_operator_groups(1)
    assert torch.equal(AS, AS_ref)
    assert dd.iter_fin == ref.iter_fin and dd.err_history == ref.err_history
    assert torch.equal(dd.X_device, ref.X_device)


def test_pass_profile_counts_and_leaves_results_alone():
    """bdda_profile_enable / bdda_profile_read: with profiling on, every
    operator application adds one timed launch to each of its five passes
    (complex128 and complex64 counted apart), the products stay bit for bit
    what they are unprofiled, and a read clears the counts."""
    say("pass profile")
    prob = _problem(L=9)
    dd = _gpu_dd(prob, "bicgstab", 1e-2, 1)
    plan = dd._plan
    g = torch.Generator(device="cuda").manual_seed(3)
    S = torch.randn(dd._B_dev.shape, dtype=torch.complex128, device="cuda", generator=g)
    plain = plan.block_mvp(S)
    plan.profile_read()
    plan.profile_enable(True)
    try:
        outs = [plan.block_mvp(S) for _ in range(3)]
        prof = plan.profile_read()
        again = plan.profile_read()
    finally:
        plan.profile_enable(False)
    assert all(torch.equal(o, plain) for o in outs)
    assert set(prof) == {("fp64", p) for p in ("z_fwd", "y_fwd", "x_conv", "y_inv", "z_inv")}
    assert all(cnt == 3 and ms > 0 for ms, cnt in prof.values())
    assert again == {}
    plan.block_mvp(S)
    assert plan.profile_read() == {}


def test_operator_groups_out_of_range_refused():
    from bdda_amd import _lib
    lib = _lib.load()
    assert lib.bdda_set_operator_groups(0) == _lib.E_ARG
    assert lib.bdda_set_operator_groups(5) == _lib.E_ARG
