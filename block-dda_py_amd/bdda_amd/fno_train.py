"""Training data from the solver, the training loop, and warm-start use.

Data: random GRE particles (bc, ab in [1, 1.5], beta in [0, 0.2], r_v giving
a lattice that fits the G^3 cube) with per-axis m_p (real part in
[1.3, 1.7], imaginary part log-uniform in [1e-3, 1e-1]), m_m = 1, 10-20
dipoles per wavelength and L random orientations, each solved on the device
(mixed precision, block-BiCGSTAB to 1e-5).  The correlation matrix of the
GRE log-radius field depends only on (bc, ab), so its eigendecomposition is
computed once per shape class and reused with fresh normal draws.

Targets are the local fields E_loc = X / alpha on the cube; the loss is the
masked relative L2 error over occupied cells.  ``warm_start`` turns a
trained network into an initial guess X0 for ``DiscreteDipoles.X0``.
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib
from ._lib import check
from .fno import FNO, _p, _stream
from .workloads import euler_angles


class ShapeBank:
    """Cached eigendecompositions of the GRE correlation matrix per (bc, ab)."""

    def __init__(self, seed=0):
        self.rng = np.random.default_rng(seed)
        self.cache = {}

    def sample_field(self, shape):
        key = (round(shape.bc_ratio, 6), round(shape.ab_ratio, 6))
        if key not in self.cache:
            w, u = np.linalg.eigh(shape.correlation_matrix())
            self.cache[key] = u * np.sqrt(np.clip(w, 0.0, None))[None, :]
        z = self.rng.standard_normal(shape.N_THETA * shape.N_PHI)
        return (shape.beta * (self.cache[key] @ z)).reshape(shape.N_THETA, shape.N_PHI)


def cube_geometry(dd, G):
    """Occupied cell indices of dd on the G^3 cube, occupancy bytes, origin / lf."""
    nx, ny, nz = (int(v) for v in dd.lattice_n)
    if max(nx, ny, nz) > G:
        raise ValueError(f"lattice {nx, ny, nz} does not fit a {G}^3 cube")
    addr = dd.lattice_address_in_target.astype(np.int64)
    ix, rem = np.divmod(addr, ny * nz)
    iy, iz = np.divmod(rem, nz)
    cells = (ix * G + iy) * G + iz
    occ = np.zeros(G ** 3, dtype=np.uint8)
    occ[cells] = 1
    origin = dd.lattice_grid_points[0] / dd.lattice_lf
    return cells, occ, origin


def sample_desc(dd, origin):
    """(L, 20) float32 per-orientation descriptors for bdda_fno_inputs."""
    chi = dd.eper_r[0] - 1.0
    kl = dd.k * dd.lattice_lf
    L = dd.L
    d = np.zeros((L, 20), dtype=np.float32)
    d[:, 0:6] = np.stack([chi.real, chi.imag], -1).reshape(6)[None, :]
    d[:, 6:9] = dd.u_inc_vec_P
    e0 = dd.e0_inc_vec_P
    d[:, 9:15] = np.stack([e0.real, e0.imag], -1).reshape(L, 6)
    d[:, 15] = kl
    d[:, 16:19] = origin[None, :]
    return d


def build_inputs(lib, occ_dev, inst_dev, desc_dev, G, n):
    x = torch.empty((n, 14, G ** 3), dtype=torch.float32, device="cuda")
    check(lib.bdda_fno_inputs(_p(x), ctypes.c_void_p(occ_dev.data_ptr()), ctypes.c_void_p(inst_dev.data_ptr()),
                              _p(desc_dev), G, n, _stream()), "fno_inputs")
    return x


def random_dipoles(rng, bank, G, L):
    """One random particle that fits a G^3 cube, as a DiscreteDipoles with B
    on the device.  r_v is drawn from [0.12, 0.45] (G / 32)^2: the lattice's
    cells per axis grow as sqrt(r_v) (the spacing grows as sqrt(c)), so the
    particles fill the cube alike at every G (the training set is G = 32)."""
    from bl_dda.scatterer import DiscreteDipoles, IncidentField, Target
    from shape_model.gaussian_ellipsoid import gaussian_ellipsoid_shape_model as Shape
    DiscreteDipoles.verbose_default = False
    for _ in range(50):
        bc, ab = rng.choice([1.0, 1.25, 1.5]), rng.choice([1.0, 1.25, 1.5])
        beta = rng.uniform(0.0, 0.2)
        r_v = rng.uniform(0.12, 0.45) * (G / 32.0) ** 2
        sh = Shape(r_v, bc, ab, beta)
        pts, _ = sh.surface_from_sample(bank.sample_field(sh))
        _, n, grid = sh.create_cuboid_lattice_that_encloses_GRE_shape(pts)
        if n.max() > G:
            continue
        dist = sh.find_nearest_distance_from_the_GRE_surf(grid, pts)
        is_in = sh.extract_lattice_address_in_GRE_volume(sh.lattice_lf, sh.distance_factor, n, dist)
        if is_in.sum() < 20:
            continue
        m_p = rng.uniform(1.3, 1.7, 3) + 1j * np.exp(rng.uniform(np.log(1e-3), np.log(1e-1), 3))
        dpl = rng.uniform(10.0, 20.0)
        wl = dpl * sh.lattice_lf * np.abs(np.max(m_p))
        tg = Target("train", n, sh.lattice_lf, grid, is_in, m_p)
        inc = IncidentField(wl, 1.0, euler_angles(L, int(rng.integers(1 << 30))))
        dd = DiscreteDipoles(tg, inc)
        dd.verbose = False
        dd.set_interaction_matrix()
        return dd
    raise RuntimeError("could not draw a particle that fits the cube")


def generate(n_inst, G=32, L=8, seed=0, tol=1e-5, log=print):
    """Solve n_inst random particles; returns device tensors for training."""
    lib = _lib.load()
    rng = np.random.default_rng(seed)
    bank = ShapeBank(seed + 1)
    S = G ** 3
    occ = torch.zeros((n_inst, S), dtype=torch.uint8, device="cuda")
    target = torch.zeros((n_inst * L, 6, S), dtype=torch.float32, device="cuda")
    desc = np.zeros((n_inst * L, 20), dtype=np.float32)
    iters = []
    t0 = time.time()
    for i in range(n_inst):
        dd = random_dipoles(rng, bank, G, L)
        dd.precision, dd.solver, dd.tol, dd.itermax = "mixed", "bicgstab", tol, 200
        dd.solve_matrix_equation()
        iters.append(dd.iter_fin + 1)
        cells, occ_h, origin = cube_geometry(dd, G)
        occ[i] = torch.from_numpy(occ_h).cuda()
        desc[i * L:(i + 1) * L] = sample_desc(dd, origin)
        cells_d = torch.from_numpy(cells).cuda()
        alpha = torch.from_numpy(np.ascontiguousarray(dd.alpha_E[0], dtype=np.complex128)).cuda()
        check(lib.bdda_cube_exchange(_p(target, i * L * 6 * S), ctypes.c_void_p(dd.X_device.data_ptr()),
                                     ctypes.c_void_p(cells_d.data_ptr()), ctypes.c_void_p(alpha.data_ptr()),
                                     len(cells), L, G, 1, _stream()), "cube_exchange")
        if log and (i + 1) % max(1, n_inst // 10) == 0:
            log(f"generated {i + 1}/{n_inst} particles ({time.time() - t0:.0f} s, mean iterations "
                f"{np.mean(iters):.1f})")
    inst = torch.from_numpy(np.repeat(np.arange(n_inst, dtype=np.int32), L)).cuda()
    return dict(occ=occ, target=target, desc=torch.from_numpy(desc).cuda(), inst=inst, G=G, L=L,
                iters=np.array(iters))


def evaluate(model, data, idx, batch=8):
    lib = _lib.load()
    G, S = data["G"], data["G"] ** 3
    losses = []
    for s in range(0, len(idx), batch):
        b = torch.as_tensor(idx[s:s + batch], device="cuda", dtype=torch.long)
        n = len(b)
        inst = data["inst"][b].contiguous()
        x = build_inputs(lib, data["occ"], inst, data["desc"][b].contiguous(), G, n)
        y = model.forward(x)
        loss = torch.empty(n, dtype=torch.float32, device="cuda")
        t = data["target"][b].contiguous()
        check(lib.bdda_masked_l2(_p(y), _p(t), ctypes.c_void_p(data["occ"].data_ptr()),
                                 ctypes.c_void_p(inst.data_ptr()), 6, S, _p(loss), None, n, _stream()), "masked_l2")
        losses.append(loss.cpu().numpy())
    return float(np.concatenate(losses).mean())


def train(model, data, steps=2000, batch=8, lr=2e-3, val_frac=0.05, seed=0, log=print):
    """Adam on the masked relative L2 loss; cosine learning-rate decay."""
    lib = _lib.load()
    G, S = data["G"], data["G"] ** 3
    n = data["target"].shape[0]
    n_inst = data["occ"].shape[0]
    n_val_inst = max(1, int(val_frac * n_inst))
    L = data["L"]
    val = np.arange(n - n_val_inst * L, n)
    tr = np.arange(0, n - n_val_inst * L)
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.time()
    for step in range(steps):
        b = torch.as_tensor(rng.choice(tr, batch, replace=False), device="cuda", dtype=torch.long)
        inst = data["inst"][b].contiguous()
        x = build_inputs(lib, data["occ"], inst, data["desc"][b].contiguous(), G, batch)
        y = model.forward(x, keep=True)
        loss = torch.empty(batch, dtype=torch.float32, device="cuda")
        dy = torch.empty_like(y)
        t = data["target"][b].contiguous()
        check(lib.bdda_masked_l2(_p(y), _p(t), ctypes.c_void_p(data["occ"].data_ptr()),
                                 ctypes.c_void_p(inst.data_ptr()), 6, S, _p(loss), _p(dy), batch, _stream()),
              "masked_l2")
        cur_lr = lr * 0.5 * (1 + np.cos(np.pi * step / steps))
        model.adam_step(model.backward(dy), lr=cur_lr)
        if step % 50 == 0 or step == steps - 1:
            hist.append((step, float(loss.mean())))
        if log and (step % max(1, steps // 10) == 0 or step == steps - 1):
            log(f"step {step}/{steps} train loss {float(loss.mean()):.4f} ({time.time() - t0:.0f} s)")
    return dict(history=hist, val_loss=evaluate(model, data, val), val_index=val, train_time=time.time() - t0)


def warm_start(model, dd, project=True, chunk=64, residuals=None):
    """Initial guess X0 (3N, L) complex128 on the device for a set-up
    DiscreteDipoles, predicted on a cube of the model's grid (the lattice is
    padded into it), `chunk` orientations per forward pass.  With `project`
    the guess is replaced by its minimal-residual combination X0 C
    (Plan.project_guess); `residuals`, if a list, receives the relative
    residuals before and after."""
    lib = _lib.load()
    G = model.G
    cells, occ_h, origin = cube_geometry(dd, G)
    occ = torch.from_numpy(occ_h).cuda()
    desc = torch.from_numpy(sample_desc(dd, origin)).cuda()
    L = dd.L
    X0 = torch.zeros((3 * dd.num_element_occupy, L), dtype=torch.complex128, device="cuda")
    cells_d = torch.from_numpy(cells).cuda()
    alpha = torch.from_numpy(np.ascontiguousarray(dd.alpha_E[0], dtype=np.complex128)).cuda()
    for l0 in range(0, L, chunk):
        n = min(chunk, L - l0)
        inst = torch.zeros(n, dtype=torch.int32, device="cuda")
        x = build_inputs(lib, occ, inst, desc[l0:l0 + n].contiguous(), G, n)
        y = model.forward(x)
        part = X0 if n == L else torch.zeros((3 * dd.num_element_occupy, n), dtype=torch.complex128, device="cuda")
        check(lib.bdda_cube_exchange(_p(y), ctypes.c_void_p(part.data_ptr()), ctypes.c_void_p(cells_d.data_ptr()),
                                     ctypes.c_void_p(alpha.data_ptr()), len(cells), n, G, 0, _stream()),
              "cube_exchange")
        if part is not X0:
            X0[:, l0:l0 + n] = part
    if project:
        r = dd._plan.project_guess(dd._B_dev, X0)
        if residuals is not None:
            residuals.extend(r)
    return X0


def at_grid(model, G):
    """The same trained weights evaluated on a G^3 cube (the spectral layers
    keep the same mode indices)."""
    return FNO(G=G, width=model.C, modes=model.K, layers=model.nl, cin=model.cin, cout=model.cout,
               hidden=model.hid, params=model.state())
