#!/usr/bin/env python3
"""Benchmark of the block-DDA solve on one MI355X.

Workload (one "step" = one full solve): GRE particle bc = 0.8, ab = 0.7,
beta = 0.2 (seed 0, its recorded coarse field) on a 51 x 48 x 63 lattice
(21,113 dipoles) padded to a 64^3 cuboid (128^3 doubled FFT grid),
birefringent m_p = (1.55, 1.60, 1.50) + 0.001i, m_m = 1.33, L = 64
orientations, complex128 block-BiCGSTAB to relative residual 1e-5 with B
resident on the device.  Every timed float64 solve recomputes its true
residual (one more operator application) after stopping on its recursive
one, as the reference's solver stops; a mixed solve checks its true
residual as part of the method.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-mixed] [--no-warm] [--no-trim] [--no-large]
                    [--no-batch] [--no-cpu-baseline] [--groups]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "block-dda_py_amd"))

METRIC = "orientations solved/sec at rel. residual 1e-5 (mixed+warm vs plain FP64), MVP HBM GB/s"
HBM_PEAK_GBS = 8000.0
PASSES = ["z_fwd", "y_fwd", "x_conv", "y_inv", "z_inv"]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_calibration(timed=20):
    """A plain device copy of known size (2 GiB read, 2 GiB written, three
    times) for scaling the FETCH_SIZE / WRITE_SIZE counters of rocprofv3;
    then `timed` more copies give the rate a plain streaming kernel reaches."""
    import torch
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(timed):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / timed
    print(json.dumps({"copy_bytes_each": n * 8, "copies": 3 + timed, "ms_per_copy": round(dt * 1e3, 4),
                      "copy_GBs_read_plus_write": round(2 * n * 8 / dt / 1e9, 1)}), flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "x_conv_f64_gre64_L64_traffic.md")


def measured_traffic():
    """HBM bytes per launch of the dominant kernel as measured with rocprofv3
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy calibration), read from
    its committed summary; None when there is none."""
    import re
    if not os.path.exists(TRAFFIC_FILE):
        return None
    m = re.search(r"traffic per launch:\s*([0-9]+)\s*bytes", open(TRAFFIC_FILE).read())
    return int(m.group(1)) if m else None


def pass_bytes(dims, m, N, L, Lb, s, extent=None, lines=None):
    """Algorithmic HBM bytes of each FFT pass for one product (each pass reads
    and writes its pruned array once; the x pass also reads the 6-component
    half spectrum once).  With `extent` (the occupied extent ox, oy, oz) and
    `lines` (the z lines holding an occupied cell) the count covers what the
    passes process: x rows below ox, y rows below oy, and only the occupied
    z lines in the z passes and the y inverse pass's stores; without them,
    the whole cuboid."""
    nx, ny, nz = extent if extent is not None else dims
    mx, my, mz = m
    Lp = -(-L // Lb) * Lb
    c = 3 * Lp
    t1 = nx * ny * mz * c
    t1_occ = (lines if lines is not None else nx * ny) * mz * c
    t2 = nx * my * mz * c
    half = (mx // 2 + 1) * (my // 2 + 1) * (mz // 2 + 1) * 6
    occ = 3 * N * L
    return {
        "z_fwd": s * (occ + t1_occ),
        "y_fwd": s * (t1 + t2),
        "x_conv": s * (2 * t2 + half),
        "y_inv": s * (t2 + t1_occ),
        "z_inv": s * (t1_occ + 2 * occ) + 2 * 16 * 3 * N,
    }


def timed_solves(dd, precision, steps, warmup, torch, profile_steps=1):
    """Warm-up, then `steps` timed solves with nothing else recorded; then
    `profile_steps` further solves with per-pass HIP events on the solve's
    stream for the pass timings (events are kept out of the timed solves).
    Every solve ends with its true residual computed."""
    dd.precision = precision
    dd.check_true_residual = True
    for _ in range(warmup):
        dd.solve_matrix_equation()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dd.solve_matrix_equation()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    dd._plan.profile_read()  # discard anything recorded so far
    dd._plan.profile_enable(True)
    for _ in range(profile_steps):
        dd.solve_matrix_equation()
    prof = dd._plan.profile_read()
    dd._plan.profile_enable(False)
    return dt, prof


_CPU_SAMPLE = r"""
import json, os, sys, time
import numpy as np
sys.path.insert(0, sys.argv[2])
from oracle import bl_krylov as okr, scatterer as osc
from oracle.mvp_fft import fft_init
d = np.load(sys.argv[1])
tg = osc.Target(d["n"], float(d["lf"]), d["grid"], d["is_in"], d["m_p"])
inc = osc.Incident(float(d["wl0"]), float(d["m_m"]), d["euler"])
eps_r, alpha, diag = osc.polarizability(tg, inc)
At = fft_init(tg.lattice_n, 3, tg.lf, inc.k)
E_inc, B = osc.rhs(tg, inc)
iters = int(sys.argv[3])
t0 = time.perf_counter()
okr.bl_bicgstab(tg.lattice_n, 3, tg.address, At, diag, B, 0.0, iters)
print(json.dumps({"seconds": time.perf_counter() - t0}))
"""


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(geometry, dd, wl, gpu_iters, max_iters=6):
    """The float64 NumPy/SciPy oracle (the reference's algorithm: scipy.fft,
    numpy matmul and LAPACK solves in its order) on the host cores: a child
    process with the FFT workers and BLAS threads set to the cores it may
    use (the box's CPU share: 16), timing min(max_iters, GPU iterations)
    block-BiCGSTAB iterations at the full L on the particle's own
    (unpadded) lattice -- the whole solve when the GPU needs no more --
    scaled to the GPU's float64 iteration count."""
    sample_iters = min(max_iters, gpu_iters)
    import subprocess
    import tempfile
    import numpy as np
    cores = min(16, len(os.sched_getaffinity(0)))
    name, n, lf, grid, is_in = geometry
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "sample.npz")
        np.savez(path, n=np.asarray(n), lf=lf, grid=grid, is_in=is_in,
                 m_p=np.array(wl["m_p_xyz"], dtype=np.complex128), wl0=dd.wl_0, m_m=wl["m_m"],
                 euler=np.asarray(dd.euler_angles))
        env = dict(os.environ, DDA_FFT_WORKERS=str(cores))
        for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
            env.setdefault(k, str(cores))
        r = subprocess.run([sys.executable, "-c", _CPU_SAMPLE, path, ROOT, str(sample_iters)], env=env,
                           capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline failed: " + r.stderr[-2000:])
    t = json.loads(r.stdout.strip().splitlines()[-1])["seconds"]
    per_iter = t / sample_iters
    total = per_iter * gpu_iters
    return {"value": round(wl["L"] / total, 4), "unit": "orientations/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "threads": {"DDA_FFT_WORKERS": cores,
                                                  "OMP_NUM_THREADS": env["OMP_NUM_THREADS"],
                                                  "OPENBLAS_NUM_THREADS": env["OPENBLAS_NUM_THREADS"]},
            "sample": f"{sample_iters} block-BiCGSTAB iterations of the float64 NumPy/SciPy oracle at L={wl['L']} "
                      f"on the particle's own {tuple(int(v) for v in n)} lattice ({t:.1f} s)"
                      + ("" if sample_iters == gpu_iters else
                         f", extrapolated to {gpu_iters} iterations (the GPU float64 count; same algorithm)")
                      + ", without a true-residual check"}


def warm_leg(dd, steps, torch, t_fp64, S_fw64, S_bk64):
    """Mixed precision with the learned warm start (the network trained on
    32^3 particles, run on the 64^3 cube): each step runs the network's
    forward pass for all L orientations, the guess's minimal-residual
    projection (one float64 operator application) and the mixed solve, whose
    rule keeps the guess only if its residual is below that of zero."""
    import numpy as np
    from bdda_amd.fno import FNO
    from bdda_amd.fno_train import warm_start
    model = FNO.load(os.path.join(ROOT, "block-dda_py_amd", "bdda_amd", "data", "fno_gre32.npz")).at_grid(64)
    dd.precision = "mixed"
    res = []

    def step():
        dd.X0 = warm_start(model, dd, residuals=res)
        dd.solve_matrix_equation()

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res.clear()
        step()
    torch.cuda.synchronize()
    dtw = (time.perf_counter() - t0) / steps
    S_fw = np.stack(dd.compute_PCAS_observable_S_fw())
    S_bk = dd.compute_OCBS_observable_S_bk()
    out = {"value": round(dd.L / dtw, 3), "ms_per_step": round(dtw * 1e3, 2), "iterations": dd.iter_fin + 1,
           "guess_used": bool(dd.warm_start_used), "guess_residual_raw": res[0], "guess_residual_projected": res[1],
           "true_residual": dd.true_residual, "speedup_vs_fp64": round(t_fp64 / dtw, 3),
           "S_fw_err_vs_fp64_of_max": float(np.max(np.abs(S_fw - S_fw64)) / np.max(np.abs(S_fw64))),
           "S_bk_err_vs_fp64_of_max": float(np.max(np.abs(S_bk - S_bk64)) / np.max(np.abs(S_bk64))),
           "network": "fno_gre32 (4 Fourier layers, width 32, 12 modes) at grid 64"}
    dd.X0 = None
    return out


def groups_leg(dd, steps, torch, t_fp64, t_mixed, groups=4):
    """The same float64 and mixed solves with the operator's column tiles in
    `groups` groups on staggered streams (bdda_set_operator_groups): one
    group's x pass overlaps the next group's z and y passes.  The results
    are bit-identical to the plain operator's."""
    from bdda_amd import _lib
    lib = _lib.load()
    X_ref = dd.X_device.clone()
    out = {"groups": groups, "note": "opt-in (bdda_set_operator_groups); the headline value and the roofline "
                                     "use the plain operator"}
    _lib.check(lib.bdda_set_operator_groups(groups))
    try:
        for prec, t_ref in (("fp64", t_fp64), ("mixed", t_mixed)):
            dtg, _ = timed_solves(dd, prec, steps, 1, torch, profile_steps=0)
            out[prec] = {"ms_per_step": round(dtg / steps * 1e3, 2), "speedup_vs_plain": round(t_ref / (dtg / steps), 3)}
            if prec == "mixed":
                out[prec]["identical_to_plain"] = bool(torch.equal(dd.X_device, X_ref))
    finally:
        lib.bdda_set_operator_groups(1)
    return out


def trimmed_leg(wl, geometry, steps, torch, t_fp64, S_fw64, S_bk64):
    """The benchmark particle without the 64^3 padding: the solver's cuboid
    is then the smallest box of lattice cells holding the occupied sites
    (what a DiscreteDipoles without pad_to computes on), with its own FFT
    lengths.  Float64 and mixed solves, their S against the padded float64
    solve's."""
    import numpy as np
    from bdda_amd import workloads
    dd = workloads.make_dipoles(dict(wl, pad_to=None), geometry=geometry, precision="fp64")
    out = {"cuboid": list(dd._plan.n), "fft_grid": list(dd._plan.m), "unit": "orientations/s"}
    for prec in ("fp64", "mixed"):
        dt, _ = timed_solves(dd, prec, steps, 1, torch, profile_steps=0)
        S_fw = np.stack(dd.compute_PCAS_observable_S_fw())
        S_bk = dd.compute_OCBS_observable_S_bk()
        out[prec] = {"value": round(dd.L * steps / dt, 3), "ms_per_step": round(dt / steps * 1e3, 2),
                     "iterations": dd.iter_fin + 1, "true_residual": dd.true_residual,
                     "speedup_vs_padded_fp64": round(t_fp64 / (dt / steps), 3),
                     "S_fw_err_vs_padded_fp64_of_max": float(np.max(np.abs(S_fw - S_fw64)) / np.max(np.abs(S_fw64))),
                     "S_bk_err_vs_padded_fp64_of_max": float(np.max(np.abs(S_bk - S_bk64)) / np.max(np.abs(S_bk64)))}
    return out


def large_leg(torch, steps=2):
    """The 96^3 configuration (config-1 physics at four times the size,
    L = 32, block-COCG-rq to 1e-5): float64 against mixed precision, with
    the S errors of the mixed solve relative to the float64 one."""
    import numpy as np
    from bdda_amd import workloads
    wl = workloads.LARGE_96
    dd = workloads.make_dipoles(wl, precision="fp64")
    out = {"workload": wl["name"], "lattice": [int(v) for v in dd.lattice_n], "cuboid": list(dd._plan.n),
           "fft_grid": list(dd._plan.m), "dipoles": int(dd.num_element_occupy), "L": wl["L"],
           "solver": wl["solver"], "tol": wl["tol"], "unit": "orientations/s"}
    S = {}
    for prec in ("fp64", "mixed"):
        dt, _ = timed_solves(dd, prec, steps, 1, torch, profile_steps=0)
        S[prec] = (np.stack(dd.compute_PCAS_observable_S_fw()), dd.compute_OCBS_observable_S_bk())
        out[prec] = {"value": round(wl["L"] * steps / dt, 2), "ms_per_solve": round(dt / steps * 1e3, 2),
                     "iterations": dd.iter_fin + 1, "true_residual": dd.true_residual}
        if prec == "mixed":
            out[prec]["refinement"] = dd.solve_info
    out["mixed"]["speedup_vs_fp64"] = round(out["mixed"]["value"] / out["fp64"]["value"], 3)
    out["mixed"]["S_fw_max_rel_err_vs_fp64"] = float(np.max(np.abs(S["mixed"][0] - S["fp64"][0]) / np.abs(S["fp64"][0])))
    out["mixed"]["S_bk_max_rel_err_vs_fp64"] = float(np.max(np.abs(S["mixed"][1] - S["fp64"][1]) / np.abs(S["fp64"][1])))
    return out


def batch_leg(torch, nstreams=8):
    """Batched uneven instances (32 GRE particles, lattices 20..48 cells,
    L = 16, block-COCG-rq to 1e-5): per precision, one-at-a-time solves
    against one concurrent batch on `nstreams` HIP streams (both timed after
    a warm-up)."""
    import numpy as np
    from bdda_amd import workloads
    from bdda_amd.batch import autotune, solve_batch
    wl = workloads.BATCH_32
    insts = workloads.batch_instances(wl)
    t0 = time.time()
    geos = [workloads.build_geometry(w) for w in insts]
    t_geo = time.time() - t0
    dds = [workloads.make_dipoles(w, geometry=g, precision=wl["precision"]) for w, g in zip(insts, geos)]
    n_or = sum(dd.L for dd in dds)
    N = [int(dd.num_element_occupy) for dd in dds]
    out = {"workload": wl["name"], "instances": len(dds), "unit": "orientations/s", "orientations": n_or,
           "solver": wl["solver"], "nstreams": nstreams, "dipoles": {"min": min(N), "max": max(N), "total": sum(N)},
           "longest_axis": {"min": int(min(max(dd.lattice_n) for dd in dds)),
                            "max": int(max(max(dd.lattice_n) for dd in dds))},
           "geometry_s": round(t_geo, 1)}
    for prec in ("mixed", "fp64"):
        for dd in dds:
            dd.precision = prec
            dd.solve_matrix_equation()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for dd in dds:
            dd.solve_matrix_equation()
        torch.cuda.synchronize()
        t_seq = time.perf_counter() - t0
        X_seq = [dd.X_device.clone() for dd in dds]
        solve_batch(dds, nstreams)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solve_batch(dds, nstreams)
        torch.cuda.synchronize()
        t_bat = time.perf_counter() - t0
        its = [dd.iter_fin + 1 for dd in dds]
        out[prec] = {
            "value": round(n_or / t_bat, 2), "batched_s": round(t_bat, 4), "sequential_s": round(t_seq, 4),
            "sequential_value": round(n_or / t_seq, 2), "speedup_vs_sequential": round(t_seq / t_bat, 3),
            "identical_to_sequential": all(torch.equal(dd.X_device, x) for dd, x in zip(dds, X_seq)),
            "all_converged": all(dd.converge for dd in dds),
            "max_true_residual": max(dd.true_residual for dd in dds),
            "iterations": {"min": min(its), "max": max(its), "mean": round(float(np.mean(its)), 2)},
        }
        # the same batch with every plan's column tile and x pass kernel
        # autotuned (tuning is plan set-up, outside the timed region)
        t0 = time.perf_counter()
        picks = autotune(dds)
        tiles, xks = [t for t, _ in picks], [k for _, k in picks]
        torch.cuda.synchronize()
        t_tune = time.perf_counter() - t0
        X_bat = [dd.X_device.clone() for dd in dds]
        solve_batch(dds, nstreams)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solve_batch(dds, nstreams)
        torch.cuda.synchronize()
        t_tuned = time.perf_counter() - t0
        out[prec]["autotuned"] = {
            "value": round(n_or / t_tuned, 2), "batched_s": round(t_tuned, 4), "tuning_s": round(t_tune, 3),
            "tiles": {str(t): tiles.count(t) for t in sorted(set(tiles))},
            "x_kernels": {{1: "split", 2: "one-pass"}[k]: xks.count(k) for k in sorted(set(xks))},
            "max_rel_diff_vs_untuned": max(float((dd.X_device - x).abs().max() / x.abs().max())
                                           for dd, x in zip(dds, X_bat)),
            "all_converged": all(dd.converge for dd in dds)}
        for dd in dds:
            dd._plan.set_column_tile(0)
            dd._plan.set_x_kernel(0)
    out["value"] = out[wl["precision"]]["value"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-warm", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-trim", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--groups", action="store_true",
                    help="also time the solves with the operator's column tiles in 4 groups on staggered streams "
                         "(kept out of the default run so that every launch of a pass covers all column tiles)")
    ap.add_argument("--copy-calibration", action="store_true", help="only run the copy used to scale HBM counters")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the multi-GPU layout is a design note, "
              "not a measured path", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from bdda_amd import workloads
    from bl_dda.scatterer import DiscreteDipoles
    DiscreteDipoles.verbose_default = False
    wl = workloads.FULL_64
    log("building geometry")
    geometry = workloads.build_geometry(wl)
    log("setting interaction matrix")
    dd = workloads.make_dipoles(wl, geometry=geometry, precision="fp64")
    plan = dd._plan
    N, L = dd.num_element_occupy, wl["L"]
    dims, m = plan.n, plan.m
    log(f"lattice {tuple(int(v) for v in dd.lattice_n)} cuboid {dims} fft {m} N={N} L={L}")

    log(f"fp64 solves: warmup {a.warmup}, steps {a.steps}")
    dt, prof = timed_solves(dd, "fp64", a.steps, a.warmup, torch)
    iters = dd.iter_fin + 1
    tr = dd.true_residual
    log(f"fp64: {dt / a.steps * 1e3:.1f} ms/solve with its true-residual check, {iters} iterations, "
        f"true residual {tr:.2e}")

    Lb = plan.column_tile(L)
    ext, nlines = dd._plan.occupied_extent, dd._plan.occupied_lines
    pb = pass_bytes(dims, m, N, L, Lb, 16, ext, nlines)
    passes = {}
    for nm in PASSES:
        ms, cnt = prof.get(("fp64", nm), (0.0, 0))
        if cnt:
            avg = ms / cnt
            passes[nm] = {"avg_ms": avg, "launches": cnt, "bytes": pb[nm], "GBs": pb[nm] / avg / 1e6}
    dom = max(passes, key=lambda k: passes[k]["avg_ms"] * passes[k]["launches"])
    d = passes[dom]
    n_cells = int(np.prod(dims))
    M = int(np.prod(m))
    mvp_alg = 16 * (26 * n_cells * 3 * L + 6 * M)
    mvp_ms = sum(p["avg_ms"] for p in passes.values())
    mvp_done = sum(p["bytes"] for p in passes.values())
    kernel = {"x_conv": "x pass (forward x FFT, 3x3 Green multiply, inverse x FFT; on m = 128 the split-spectrum "
                        "kernel k_xconv_split)", "y_fwd": "y forward pass",
              "y_inv": "y inverse pass", "z_fwd": "z forward pass (scatter)",
              "z_inv": "z inverse pass (gather)"}[dom]
    roofline = {"bound": "hbm",
                "kernel": f"{kernel}, m = {m[0]}, complex128, {Lb}-column tile, x rows below the occupied "
                          f"extent ({ext[0]} of {dims[0]})",
                "achieved": round(d["GBs"], 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(d["GBs"] / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(), "bytes_per_launch": d["bytes"],
                "avg_launch_ms": round(d["avg_ms"], 4)}

    result = {
        "metric": METRIC, "value": round(L * a.steps / dt, 3), "unit": "orientations/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(dt / a.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl["name"], "lattice": [int(v) for v in dd.lattice_n], "cuboid": list(dims),
                   "fft_grid": list(m), "dipoles": int(N), "L": L, "solver": wl["solver"], "tol": wl["tol"],
                   "precision": "fp64"},
        "iterations": iters, "true_residual": tr,
        "true_residual_check": "inside every timed solve: the float64 solve stops on its recursive residual "
                               "like the reference's, then recomputes ||jpre(B - A X)|| / ||jpre B|| (one "
                               "complex128 operator application)",
        "roofline": roofline,
        "mvp": {"ms": round(mvp_ms, 4), "alg_bytes": mvp_alg, "alg_GBs": round(mvp_alg / mvp_ms / 1e6, 1),
                "alg_bytes_note": "16 (26 n 3L + 6 M) over the whole cuboid; the passes skip the lines and rows "
                                  "beyond the target's occupied extent and the z lines with no occupied cell, "
                                  "so they move bytes_processed",
                "occupied_extent": list(ext), "occupied_z_lines": nlines, "bytes_processed": mvp_done,
                "GBs_processed": round(mvp_done / mvp_ms / 1e6, 1),
                "passes": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                           for k, v in passes.items()}},
    }

    if not a.no_mixed:
        S_fw64 = np.stack(dd.compute_PCAS_observable_S_fw())
        S_bk64 = dd.compute_OCBS_observable_S_bk()
        log("mixed solves")
        dtm, profm = timed_solves(dd, "mixed", a.steps, a.warmup, torch)
        S_fwm = np.stack(dd.compute_PCAS_observable_S_fw())
        S_bkm = dd.compute_OCBS_observable_S_bk()
        pbm = pass_bytes(dims, m, N, L, Lb, 8, ext, nlines)
        mp = {}
        for nm in PASSES:
            ms, cnt = profm.get(("fp32", nm), (0.0, 0))
            if cnt:
                mp[nm] = {"avg_ms": round(ms / cnt, 4), "launches": cnt, "GBs": round(pbm[nm] / (ms / cnt) / 1e6, 1)}
        result["mixed"] = {
            "value": round(L * a.steps / dtm, 3), "ms_per_step": round(dtm / a.steps * 1e3, 2),
            "iterations": dd.iter_fin + 1, "true_residual": dd.true_residual, "refinement": dd.solve_info,
            "speedup_vs_fp64": round(dt / dtm, 3),
            "S_fw_max_rel_err_vs_fp64": float(np.max(np.abs(S_fwm - S_fw64) / np.abs(S_fw64))),
            "S_bk_max_rel_err_vs_fp64": float(np.max(np.abs(S_bkm - S_bk64) / np.abs(S_bk64))),
            "S_fw_err_vs_fp64_of_max": float(np.max(np.abs(S_fwm - S_fw64)) / np.max(np.abs(S_fw64))),
            "S_bk_err_vs_fp64_of_max": float(np.max(np.abs(S_bkm - S_bk64)) / np.max(np.abs(S_bk64))),
            "passes_c64": mp,
        }
        log(f"mixed: {dtm / a.steps * 1e3:.1f} ms/solve")

    if a.groups and not a.no_mixed:
        result["operator_groups"] = groups_leg(dd, a.steps, torch, dt / a.steps, dtm / a.steps)
        og = result["operator_groups"]
        log(f"operator in {og['groups']} column groups: fp64 {og['fp64']['ms_per_step']} ms, "
            f"mixed {og['mixed']['ms_per_step']} ms")

    if not a.no_mixed and not a.no_trim:
        log("same particle on its occupied extent")
        result["trimmed"] = trimmed_leg(wl, geometry, a.steps, torch, dt / a.steps, S_fw64, S_bk64)
        t = result["trimmed"]
        log(f"trimmed cuboid {t['cuboid']} fft {t['fft_grid']}: fp64 {t['fp64']['ms_per_step']} ms, "
            f"mixed {t['mixed']['ms_per_step']} ms")

    if not a.no_mixed and not a.no_warm:
        log("mixed + learned warm start")
        result["mixed_warm"] = warm_leg(dd, a.steps, torch, dt / a.steps, S_fw64, S_bk64)
        w = result["mixed_warm"]
        log(f"mixed + warm: {w['ms_per_step']} ms/solve, {w['iterations']} iterations, guess used {w['guess_used']}")

    if not a.no_large:
        log("96^3 configuration")
        result["large"] = large_leg(torch)
        g = result["large"]
        log(f"96^3: fp64 {g['fp64']['ms_per_solve']} ms, mixed {g['mixed']['ms_per_solve']} ms")

    if not a.no_batch:
        log("batched uneven instances")
        result["batch"] = batch_leg(torch)
        b = result["batch"]
        log(f"batch: mixed {b['mixed']['value']} / fp64 {b['fp64']['value']} orientations/s "
            f"(x{b['mixed']['speedup_vs_sequential']} / x{b['fp64']['speedup_vs_sequential']} vs one at a time)")
        log(f"batch, autotuned tiles and x kernels: mixed {b['mixed']['autotuned']['value']} / fp64 "
            f"{b['fp64']['autotuned']['value']} orientations/s (tiles {b['fp64']['autotuned']['tiles']}, "
            f"x kernels {b['fp64']['autotuned']['x_kernels']})")

    if not a.no_cpu_baseline:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(geometry, dd, wl, iters)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
