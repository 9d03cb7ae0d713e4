"""Block-Krylov solvers with Jacobi preconditioning on the GPU.

Drop-in for the reference ``bl_krylov/bl_krylov.py`` (v0.3.0): the same
``_build_scatter_index``, ``_block_mvp``, ``bl_cocg_rq_jacobi_mvp_fft`` and
``bl_bicgstab_jacobi_mvp_fft`` with the same arguments and return values
``(X, iter_fin, err_fin)``.  The whole iteration runs on the device (libbdda);
only L x L coefficients visit the host.  Keyword-only extensions:

* ``precision``: ``'fp64'`` (the reference algorithm in complex128) or
  ``'mixed'`` (complex64 operator / vector updates, float64 L x L solves and
  block QR, float64 true residual and refinement every ``refine_every``
  iterations; ``err_fin`` is then the true residual).
* ``X0``: initial guess (warm start), discarded if its preconditioned
  residual exceeds ||jpre B||.
* ``verbose``: print ``iter= k, err= ...`` lines like the reference
  (``bl_krylov.py:107,178``).
"""
import numpy as np
import torch

from mvp_fft.mvp_fft import GreenTensor


def _build_scatter_index(address, f):
    """full_idx[f*m + c] = f*address[m] + c (bl_krylov.py:5-24)."""
    address = np.asarray(address)
    return np.repeat(address, f) * f + np.tile(np.arange(f), len(address))


def _target_for(Au_til, address, DIAG_A, jpre):
    if not isinstance(Au_til, GreenTensor):
        raise TypeError("Au_til must be the handle returned by mvp_fft.fft_init")
    address = np.asarray(address, dtype=np.int64)
    diag = np.asarray(DIAG_A, dtype=np.complex128)
    jp = np.asarray(jpre, dtype=np.complex128)
    key = (address.size, hash(address.tobytes()), hash(diag.tobytes()), hash(jp.tobytes()))
    plan = Au_til.plan
    if plan.target_key != key:
        plan.set_target(address, diag, jp, key=key)
    return plan


def _block_mvp(n, f, Au_til, full_idx, DIAG_A, S, jpre):
    """AS = jpre * (DIAG_A * S + A_offdiag S) (bl_krylov.py:27-55)."""
    full_idx = np.asarray(full_idx)
    if f != 3 or full_idx.size % 3:
        raise ValueError("_block_mvp: only f=3 is supported")
    address = full_idx[0::3] // 3
    plan = _target_for(Au_til, address, DIAG_A, jpre)
    is_torch = isinstance(S, torch.Tensor)
    AS = plan.block_mvp(S)
    return AS if is_torch else AS.cpu().numpy()


def _solve(which, n, f, address, Au_til, DIAG_A, B, tol, itermax, precision, X0, refine_every, verbose):
    if f != 3:
        raise ValueError("only f=3 (electric dipoles) is supported")
    jpre = 1.0 / np.asarray(DIAG_A, dtype=np.complex128)
    plan = _target_for(Au_til, address, DIAG_A, jpre)
    is_torch = isinstance(B, torch.Tensor)
    X, it, err, tr, hist, _ = plan.solve(B, which, precision, tol, itermax, X0=X0, refine_every=refine_every,
                                         true_residual=False)
    if verbose:
        for k, e in enumerate(hist):
            print("iter= {:}, err= {:.4f}".format(k, e))
    if not hist:
        err = float('inf')
    return (X if is_torch else X.cpu().numpy()), it, err


def bl_cocg_rq_jacobi_mvp_fft(n, f, address, Au_til, DIAG_A, B, tol, itermax, *, precision="fp64", X0=None,
                              refine_every=20, verbose=True):
    """Block-COCG-rq (bl_krylov.py:58-118) on the device."""
    return _solve("cocg_rq", n, f, address, Au_til, DIAG_A, B, tol, itermax, precision, X0, refine_every,
                  verbose)


def bl_bicgstab_jacobi_mvp_fft(n, f, address, Au_til, DIAG_A, B, tol, itermax, *, precision="fp64", X0=None,
                               refine_every=20, verbose=True):
    """Block-BiCGSTAB (bl_krylov.py:121-187) on the device."""
    return _solve("bicgstab", n, f, address, Au_til, DIAG_A, B, tol, itermax, precision, X0, refine_every,
                  verbose)
