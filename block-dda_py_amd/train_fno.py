"""Train the learned warm start on solver-generated data and measure its gain.

    python block-dda_py_amd/train_fno.py --instances 2000 --orientations 8 \
        --steps 3000 --out block-dda_py_amd/bdda_amd/data/fno_gre32.npz

1. draws random GRE particles on a 32^3 cube and solves each on the GPU
   (mixed precision, block-BiCGSTAB to 1e-5) for the training targets;
2. trains the 4-layer, width-32, 12-mode Fourier neural operator with its
   own forward/backward kernels (bf16 MFMA, float32 accumulation) and Adam;
3. compares mixed-precision block-COCG-rq on held-out isotropic particles
   (L = 64) from a zero start, from the network's raw guess and from its
   minimal-residual projection: iterations and time to relative residual
   1e-5, inference and projection time included.  `--eval-only --eval-grid
   64` runs the trained weights on particles that fill a 64^3 cube.
Prints one JSON summary line.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def log(msg):
    print(f"[train_fno] {msg}", file=sys.stderr, flush=True)


def gain(model, n_eval=4, L=64, G=32, seed=12345):
    """Zero start vs warm start on held-out isotropic particles (COCG-rq, mixed)."""
    import numpy as np
    import torch
    from bdda_amd.fno_train import ShapeBank, random_dipoles, warm_start
    rng = np.random.default_rng(seed)
    bank = ShapeBank(seed + 1)
    rows = []
    for i in range(n_eval):
        dd = random_dipoles(rng, bank, G, L)
        iso = dd.m_p_xyz[0]
        from bl_dda.scatterer import DiscreteDipoles, IncidentField, Target
        tg = Target("eval", dd.lattice_n, dd.lattice_lf, dd.lattice_grid_points,
                    dd.lattice_grid_points_is_in_target, np.array([iso] * 3))
        inc = IncidentField(dd.wl_0, dd.m_m, dd.euler_angles)
        dd = DiscreteDipoles(tg, inc)
        dd.verbose = False
        dd.set_interaction_matrix()
        dd.solver, dd.precision, dd.tol, dd.itermax = "cocg_rq", "mixed", 1e-5, 300
        dd.solve_matrix_equation()  # warm-up of this shape's buffers
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dd.X0 = None
        dd.solve_matrix_equation()
        torch.cuda.synchronize()
        t_zero, it_zero, res_zero = time.perf_counter() - t0, dd.iter_fin + 1, dd.true_residual
        t0 = time.perf_counter()
        dd.X0 = warm_start(model, dd, project=False)
        dd.solve_matrix_equation()
        torch.cuda.synchronize()
        t_raw, it_raw = time.perf_counter() - t0, dd.iter_fin + 1
        rr = []
        t0 = time.perf_counter()
        dd.X0 = warm_start(model, dd, residuals=rr)
        dd.solve_matrix_equation()
        torch.cuda.synchronize()
        t_warm, it_warm, res_warm = time.perf_counter() - t0, dd.iter_fin + 1, dd.true_residual
        rows.append(dict(N=int(dd.num_element_occupy), lattice=[int(v) for v in dd.lattice_n],
                         it_zero=it_zero, it_raw_guess=it_raw, it_warm=it_warm, t_zero_ms=round(t_zero * 1e3, 2),
                         t_raw_guess_ms=round(t_raw * 1e3, 2), t_warm_ms=round(t_warm * 1e3, 2),
                         guess_residual=rr[0], projected_residual=rr[1], res_zero=res_zero, res_warm=res_warm,
                         warm_used=bool(dd.warm_start_used)))
        log(f"eval {i}: {rows[-1]}")
    return rows


def guess_study(n_inst=2, L=64, G=32, seed=12345, levels=(0.3, 0.1, 0.03, 0.01, 1e-3)):
    """Iterations to 1e-5 from perturbed solutions: how much an initial guess
    of a given relative error can save (block-COCG-rq, float64 and mixed, on
    held-out isotropic particles).  The perturbation is random complex noise
    or a smooth per-orientation scale and phase error of the stated norm."""
    import numpy as np
    import torch
    from bdda_amd.fno_train import ShapeBank, random_dipoles
    from bl_dda.scatterer import DiscreteDipoles, IncidentField, Target
    rng = np.random.default_rng(seed)
    bank = ShapeBank(seed + 1)
    rows = []
    for inst in range(n_inst):
        d0 = random_dipoles(rng, bank, G, L)
        tg = Target("eval", d0.lattice_n, d0.lattice_lf, d0.lattice_grid_points,
                    d0.lattice_grid_points_is_in_target, np.array([d0.m_p_xyz[0]] * 3))
        dd = DiscreteDipoles(tg, IncidentField(d0.wl_0, d0.m_m, d0.euler_angles))
        dd.verbose = False
        dd.set_interaction_matrix()
        dd.solver, dd.precision, dd.tol, dd.itermax = "cocg_rq", "fp64", 1e-10, 300
        dd.solve_matrix_equation()
        Xs = dd.X_device.clone()
        dd.tol = 1e-5
        for prec in ("fp64", "mixed"):
            dd.precision, dd.X0 = prec, None
            dd.solve_matrix_equation()
            row = dict(instance=inst, precision=prec, zero_start=dd.iter_fin + 1)
            g = torch.Generator(device="cuda")
            g.manual_seed(0)
            for kind in ("random", "smooth"):
                for d in levels:
                    if kind == "random":
                        E = torch.randn(Xs.shape, dtype=torch.complex128, device="cuda", generator=g)
                    else:
                        c = 1 + d * torch.randn((1, Xs.shape[1]), dtype=torch.complex128, device="cuda", generator=g)
                        E = Xs * (c - 1)
                    E = E / torch.linalg.norm(E) * torch.linalg.norm(Xs) * d
                    dd.X0 = Xs + E
                    dd.solve_matrix_equation()
                    row[f"{kind}_{d:g}"] = dd.iter_fin + 1
            log(f"guess study: {row}")
            rows.append(row)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=2000)
    ap.add_argument("--orientations", type=int, default=8)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--grid", type=int, default=32)
    ap.add_argument("--eval", type=int, default=4)
    ap.add_argument("--out", default=os.path.join(HERE, "bdda_amd", "data", "fno_gre32.npz"))
    ap.add_argument("--eval-only", action="store_true", help="only measure the gain of the weights at --out")
    ap.add_argument("--eval-grid", type=int, default=0,
                    help="with --eval-only: run the network on held-out particles of this cube size (0: its own)")
    ap.add_argument("--guess-study", action="store_true", help="iterations saved by guesses of known error")
    a = ap.parse_args()
    from bdda_amd.fno import FNO
    from bdda_amd.fno_train import generate, train
    from bl_dda.scatterer import DiscreteDipoles
    DiscreteDipoles.verbose_default = False
    if a.guess_study:
        print(json.dumps(dict(guess_study=guess_study())), flush=True)
        return
    if a.eval_only:
        model = FNO.load(a.out)
        G = a.eval_grid or model.G
        if G != model.G:
            model = model.at_grid(G)
        print(json.dumps(dict(weights=a.out, grid=G, gain=gain(model, n_eval=a.eval, G=G))), flush=True)
        return
    t0 = time.time()
    data = generate(a.instances, G=a.grid, L=a.orientations, log=log)
    t_gen = time.time() - t0
    model = FNO(G=a.grid)
    res = train(model, data, steps=a.steps, batch=a.batch, lr=a.lr, log=log)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    model.save(a.out)
    rows = gain(model, n_eval=a.eval, G=a.grid)
    summary = dict(instances=a.instances, orientations=a.orientations, steps=a.steps, batch=a.batch,
                   gen_s=round(t_gen, 1), train_s=round(res["train_time"], 1),
                   mean_solver_iterations=float(data["iters"].mean()), val_loss=res["val_loss"],
                   loss_history=res["history"], gain=rows, weights=a.out)
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
