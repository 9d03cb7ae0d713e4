"""Batched uneven instances: several particles solved concurrently.

Each ``DiscreteDipoles`` keeps its own plan (FFT lengths, column tile, Green
spectrum and workspaces sized for its own lattice), so instances of different
sizes need no padding to a common grid.  ``solve_batch`` hands all of them to
``bdda_solve_batch``: host worker threads, each with its own HIP stream, take
the instances largest first and run the same device solver as
``DiscreteDipoles.solve_matrix_equation``, so a batch returns exactly the
per-instance results of one-at-a-time solves while the GPU overlaps the
instances' kernels and host round trips.
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib
from ._lib import check


def autotune(dds):
    """Fix each instance's column tile and x pass kernel to the fastest pair
    for its own lattice, L and precision (``bdda_plan_autotune``: the
    operator timed at each candidate); done once per plan, L and precision.
    Returns the (tile, x kernel) pairs."""
    picks = []
    for dd in dds:
        key = (dd._B_dev.shape[1], dd.precision)
        done = getattr(dd._plan, "_tuned", {})
        if key not in done:
            done[key] = dd._plan.autotune(*key)
            dd._plan._tuned = done
        else:
            dd._plan.set_column_tile(done[key][0])
            dd._plan.set_x_kernel(done[key][1])
        picks.append(done[key])
    return picks


def solve_batch(dds, nstreams=8):
    """Solve every set-up DiscreteDipoles in ``dds``; they must share solver,
    precision, tol, itermax and refine_every.  Per-instance results land on
    each object as after ``solve_matrix_equation``."""
    dds = list(dds)
    if not dds:
        return
    lib = _lib.load()
    first = dds[0]
    keys = ("solver", "precision", "tol", "itermax", "refine_every")
    for dd in dds:
        if dd._plan is None or dd._B_dev is None:
            raise ValueError("solve_batch: call set_interaction_matrix() on every instance first")
        for k in keys:
            if getattr(dd, k) != getattr(first, k):
                raise ValueError(f"solve_batch: instances differ in {k}")
    nb = len(dds)
    itermax = int(first.itermax)
    Xs = [torch.empty_like(dd._B_dev) for dd in dds]
    X0s = [None if dd.X0 is None else torch.as_tensor(dd.X0, dtype=torch.complex128, device="cuda")
           for dd in dds]
    for dd, x0 in zip(dds, X0s):
        if x0 is not None and tuple(x0.shape) != tuple(dd._B_dev.shape):
            raise ValueError("solve_batch: X0 must have the shape of B (3N, L)")
    vp = ctypes.c_void_p
    plans = (vp * nb)(*[dd._plan.h for dd in dds])
    Bp = (vp * nb)(*[dd._B_dev.data_ptr() for dd in dds])
    Xp = (vp * nb)(*[x.data_ptr() for x in Xs])
    X0p = (vp * nb)(*[0 if x is None else x.data_ptr() for x in X0s])
    Ls = (ctypes.c_int * nb)(*[dd._B_dev.shape[1] for dd in dds])
    it = np.zeros(nb, dtype=np.int32)
    err = np.zeros(nb, dtype=np.float64)
    tr = np.zeros(nb, dtype=np.float64)
    flags = np.zeros(nb, dtype=np.int32)
    hist = np.zeros((nb, max(itermax, 1)), dtype=np.float64)
    torch.cuda.synchronize()  # B and X0 were written on torch's stream
    t0 = time.time()
    pi = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    pd = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    check(lib.bdda_solve_batch(nb, plans, _lib.SOLVER[first.solver], _lib.PRECISION[first.precision], Bp,
                               X0p if any(x is not None for x in X0s) else None, Xp, Ls, float(first.tol), itermax,
                               int(first.refine_every), int(nstreams), pi(it), pd(err), pd(tr), pd(hist),
                               pi(flags)), "solve_batch")
    seconds = time.time() - t0
    for i, dd in enumerate(dds):
        nh = int(flags[i]) >> 8
        dd.solve_info = {"refine_cycles": (int(flags[i]) >> 2) & 63, "fallback_complex128": bool(flags[i] & 2)}
        dd._store_solution(Xs[i], int(it[i]), float(err[i]), float(tr[i]), hist[i, :nh].tolist(),
                           bool(flags[i] & 1), seconds)
    return seconds
