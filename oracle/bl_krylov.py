"""Block-Krylov solvers with Jacobi preconditioning — oracle.

Restates ``bl_krylov/bl_krylov.py`` of the reference:

* ``scatter_index``   <- ``_build_scatter_index``          (``bl_krylov.py:5-24``)
* ``block_mvp``       <- ``_block_mvp``                    (``bl_krylov.py:27-55``)
* ``bl_cocg_rq``      <- ``bl_cocg_rq_jacobi_mvp_fft``     (``bl_krylov.py:58-118``)
* ``bl_bicgstab``     <- ``bl_bicgstab_jacobi_mvp_fft``    (``bl_krylov.py:121-187``)

The operation order is the reference's: reduced Householder QR, plain
transposes in COCG-rq, conjugate transposes and the trace form of xi in
BiCGSTAB, stop after the X/R update when err < tol, 0-based ``iter_fin``.
Each solver also returns the per-iteration ``err`` history (the reference
prints it, ``bl_krylov.py:107,178``).

Pinned by: convergence of both solvers to ``numpy.linalg.solve`` of the dense
3N x 3N system (``oracle.dense``) on small targets, the QR-basis invariance of
the iterates, and end-to-end against the reference-held HDF5 results and the
``test_dda.ipynb`` recorded err sequence.
"""
import numpy as np

from .mvp_fft import fft_mvp


def scatter_index(address, f=3):
    """full_idx[f*m + c] = f*address[m] + c (``bl_krylov.py:23-24``)."""
    address = np.asarray(address, dtype=np.int64)
    return (address[:, None] * f + np.arange(f)[None, :]).ravel()


def block_mvp(n, f, A_tilde, full_idx, diag, S, jpre):
    """jpre * (diag*S + A_off S) for all columns (``bl_krylov.py:45-55``)."""
    ncell = int(np.prod(n))
    Pc = np.zeros((f * ncell, S.shape[1]), dtype=np.complex128)
    Pc[full_idx, :] = S
    AP = fft_mvp(n, f, A_tilde, Pc)
    return (diag[:, None] * S + AP[full_idx, :]) * jpre[:, None]


def bl_cocg_rq(n, f, address, A_tilde, diag, B, tol, itermax, X0=None, matvec=None):
    """Block-COCG-rq (Gu et al. 2016), ``bl_krylov.py:58-118``.

    Returns (X, iter_fin, err_fin, err_history).  ``X0`` (warm start) and
    ``matvec`` (an alternative operator S -> jpre*(A S)) are extensions used
    by tests; with the defaults the algorithm is the reference's.
    """
    jpre = 1.0 / diag
    idx = scatter_index(address, f)
    op = matvec or (lambda S: block_mvp(n, f, A_tilde, idx, diag, S, jpre))
    Bj = B * jpre[:, None]
    bnorm = np.linalg.norm(Bj)
    if X0 is None:
        X = np.zeros_like(B)
        R0 = Bj
    else:
        X = X0.copy()
        R0 = Bj - op(X0)
    Q, xi = np.linalg.qr(R0, mode='reduced')
    S = Q
    it, err_fin, hist = 0, float('inf'), []
    for k in range(itermax):
        AS = op(S)
        alpha = np.linalg.solve(S.T @ AS, Q.T @ Q)
        X = X + S @ (alpha @ xi)
        Qn, tau = np.linalg.qr(Q - AS @ alpha, mode='reduced')
        xi = tau @ xi
        err = np.linalg.norm(xi) / bnorm
        hist.append(err)
        it, err_fin = k, err
        if err < tol:
            break
        beta = np.linalg.solve(Q.T @ Q, tau.T @ (Qn.T @ Qn))
        Q = Qn
        S = Q + S @ beta
    return X, it, err_fin, hist


def bl_bicgstab(n, f, address, A_tilde, diag, B, tol, itermax, X0=None, matvec=None):
    """Block-BiCGSTAB (Tadano et al. 2009), ``bl_krylov.py:121-187``.

    Returns (X, iter_fin, err_fin, err_history).
    """
    jpre = 1.0 / diag
    idx = scatter_index(address, f)
    op = matvec or (lambda S: block_mvp(n, f, A_tilde, idx, diag, S, jpre))
    Bj = B * jpre[:, None]
    bnorm = np.linalg.norm(Bj)
    if X0 is None:
        X = np.zeros_like(B)
        R = Bj.copy()
    else:
        X = X0.copy()
        R = Bj - op(X0)
    P = R.copy()
    R0H = R.copy().conj().T
    it, err_fin, hist = 0, float('inf'), []
    for k in range(itermax):
        V = op(P)
        RV = R0H @ V
        alpha = np.linalg.solve(RV, R0H @ R)
        T = R - V @ alpha
        Z = op(T)
        xi = np.trace(Z.conj().T @ T) / np.trace(Z.conj().T @ Z)
        X = X + P @ alpha + xi * T
        R = T - xi * Z
        err = np.linalg.norm(R) / bnorm
        hist.append(err)
        it, err_fin = k, err
        if err < tol:
            break
        beta = np.linalg.solve(RV, (-R0H @ Z))
        P = R + (P - xi * V) @ beta
    return X, it, err_fin, hist


def true_residual(n, f, address, A_tilde, diag, B, X):
    """||jpre*(B - A X)||_F / ||jpre*B||_F recomputed from X (float64)."""
    jpre = 1.0 / diag
    idx = scatter_index(address, f)
    Bj = B * jpre[:, None]
    R = Bj - block_mvp(n, f, A_tilde, idx, diag, X, jpre)
    return np.linalg.norm(R) / np.linalg.norm(Bj)
