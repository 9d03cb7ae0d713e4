"""Brute-force dense DDA operator — oracle cross-check (small N only).

The Goodman FFT product of ``mvp_fft.py:122-173`` equals, exactly in real
arithmetic, the direct sum  Y_i = sum_{j != i} A'(r_i - r_j) P_j  over cuboid
cells, with A'(r) the closed form of ``mvp_fft.py:78-84``.  This module builds
that dense (3n x 3n) matrix directly from the closed form, independent of any
FFT, displacement map or zero-plane logic, to pin ``oracle.mvp_fft``.
"""
import numpy as np

from .mvp_fft import green_element


def dense_offdiag(n, lf, k, cells=None):
    """Dense off-diagonal operator over the given cuboid cells (default: all).

    Row/column ordering is 3*cell_index + c, cell index = ix*Ny*Nz+iy*Nz+iz.
    """
    nx, ny, nz = (int(v) for v in n)
    if cells is None:
        cells = np.arange(nx * ny * nz)
    ijk = np.stack(np.unravel_index(cells, (nx, ny, nz)), axis=1).astype(np.float64)
    m = len(cells)
    A = np.zeros((3 * m, 3 * m), dtype=np.complex128)
    for a in range(m):
        for b in range(m):
            if a == b:
                continue
            d = (ijk[a] - ijk[b]) * lf
            A[3 * a:3 * a + 3, 3 * b:3 * b + 3] = green_element(d[0], d[1], d[2], k)
    return A
