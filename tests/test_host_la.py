"""The host L x L algebra of the block-Krylov solvers (csrc/host_la.cpp),
built with g++ into a small driver and checked against NumPy / LAPACK: the
LU pivots equal LAPACK's zgetrf choice (numpy.linalg.solve, used by the
reference at bl_krylov.py:101,113,167,184), the solution, the Cholesky
factor, the triangular inverse and the product."""
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "block-dda_py_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("la") / "la_driver")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "host", "la_driver.cpp"),
                    os.path.join(CSRC, "host_la.cpp"), "-o", out], check=True)
    return out


def _run(driver, A, B, G):
    n, m = B.shape
    vals = [f"{n} {m}"]
    for M in (A, B, G):
        vals += [f"{float(v.real):.17g} {float(v.imag):.17g}" for v in M.ravel()]
    res = subprocess.run([driver], input="\n".join(vals) + "\n", capture_output=True, text=True, check=True)
    tok = res.stdout.split()
    pos = 0

    def take(k):
        nonlocal pos
        out = tok[pos:pos + k]
        pos += k
        return out

    def mat(rows, cols):
        v = np.array([float(t) for t in take(2 * rows * cols)])
        return (v[0::2] + 1j * v[1::2]).reshape(rows, cols)

    ok = int(take(1)[0])
    piv = np.array([int(t) for t in take(n)])
    X = mat(n, m) if ok else None
    pd = int(take(1)[0])
    R, Ri, C = (mat(n, n), mat(n, n), mat(n, n)) if pd else (None, None, None)
    return ok, piv, X, pd, R, Ri, C


@pytest.mark.parametrize("n,m,seed", [(1, 1, 0), (2, 3, 1), (7, 5, 2), (16, 16, 3), (64, 64, 4), (70, 70, 5)])
def test_host_la_matches_lapack(driver, n, m, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))
    Z = rng.standard_normal((2 * n, n)) + 1j * rng.standard_normal((2 * n, n))
    G = Z.conj().T @ Z
    ok, piv, X, pd, R, Ri, C = _run(driver, A, B, G)
    assert ok == 1 and pd == 1
    # LAPACK's pivots (0-based ipiv of zgetrf)
    _, lapack_piv = scipy.linalg.lu_factor(A)
    assert np.array_equal(piv, lapack_piv)
    X_ref = np.linalg.solve(A, B)
    assert np.linalg.norm(X - X_ref) / np.linalg.norm(X_ref) < 1e-11
    # R upper triangular with a real positive diagonal, R^H R = G
    assert np.allclose(np.tril(R, -1), 0) and np.all(np.diag(R).real > 0) and np.allclose(np.diag(R).imag, 0)
    assert np.linalg.norm(R.conj().T @ R - G) / np.linalg.norm(G) < 1e-13
    assert np.linalg.norm(R @ Ri - np.eye(n)) < 1e-10 * n
    assert np.linalg.norm(C - Ri @ Ri) / np.linalg.norm(Ri @ Ri) < 1e-13


def test_host_la_reports_singular_and_indefinite(driver):
    A = np.array([[1, 2], [2, 4]], dtype=complex)  # rank 1 with an exact zero pivot
    B = np.eye(2, dtype=complex)
    G = np.array([[1, 2], [2, 1]], dtype=complex)  # Hermitian, indefinite
    ok, _, _, pd, _, _, _ = _run(driver, A, B, G)
    assert ok == 0 and pd == 0


@pytest.fixture(scope="module")
def qr_driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("qr") / "qr_driver")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "host", "qr_driver.cpp"),
                    os.path.join(CSRC, "host_la.cpp"), "-o", out], check=True)
    return out


def _qr(qr_driver, A):
    rows, n = A.shape
    vals = [f"{rows} {n}"] + [f"{float(v.real):.17g} {float(v.imag):.17g}" for v in A.ravel()]
    res = subprocess.run([qr_driver], input="\n".join(vals) + "\n", capture_output=True, text=True, check=True)
    v = np.array([float(t) for t in res.stdout.split()])
    z = v[0::2] + 1j * v[1::2]
    return z[:rows * n].reshape(rows, n), z[rows * n:].reshape(n, n)


@pytest.mark.parametrize("rows,n,seed", [(1, 1, 0), (5, 1, 1), (40, 7, 2), (300, 24, 3), (64, 64, 4)])
def test_householder_qr_matches_numpy(qr_driver, rows, n, seed):
    """The block QR's last resort is LAPACK's Householder QR (zgeqrf, zungqr)
    as numpy.linalg.qr runs it in the reference (bl_krylov.py:72,110): the
    same reflectors and signs, so Q and R equal numpy's to rounding."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((rows, n)) + 1j * rng.standard_normal((rows, n))
    Q, R = _qr(qr_driver, A)
    Qn, Rn = np.linalg.qr(A, mode="reduced")
    assert np.allclose(np.tril(R, -1), 0) and np.allclose(np.diag(R).imag, 0)
    assert np.linalg.norm(R - Rn) / np.linalg.norm(Rn) < 1e-13
    assert np.linalg.norm(Q - Qn) / np.linalg.norm(Qn) < 1e-13
    assert np.linalg.norm(Q.conj().T @ Q - np.eye(n)) < 1e-13 * n


def test_householder_qr_of_a_rank_deficient_block(qr_driver):
    """Dependent columns (where Cholesky QR breaks down): Q still has
    orthonormal columns and Q R reproduces the block."""
    rng = np.random.default_rng(9)
    A = rng.standard_normal((60, 6)) + 1j * rng.standard_normal((60, 6))
    A[:, 3] = A[:, 0] + (2 - 1j) * A[:, 1]
    A[:, 5] = 1j * A[:, 2]
    Q, R = _qr(qr_driver, A)
    assert np.linalg.norm(Q.conj().T @ Q - np.eye(6)) < 1e-13
    assert np.linalg.norm(Q @ R - A) / np.linalg.norm(A) < 1e-14
    assert abs(R[3, 3]) < 1e-13 * np.linalg.norm(A) and abs(R[5, 5]) < 1e-13 * np.linalg.norm(A)
