"""Goodman doubled-grid FFT interaction operator — oracle.

Restates ``mvp_fft/mvp_fft.py`` of the reference:

* ``interaction_tensor``  <- ``build_interaction_tensor`` (``mvp_fft.py:13-94``)
* ``fft_init``            <- ``fft_init``                 (``mvp_fft.py:97-119``)
* ``fft_mvp``             <- ``fft_mvp``                  (``mvp_fft.py:122-173``)

Pinned by: the per-element closed form A'(r) and the exact symmetry
A'(r)[i,j] == A'(r)[j,i] (tests), and ``fft_mvp`` == the direct dense sum
sum_j A'(r_i - r_j) P_j on small grids (``oracle.dense``), which fixes the
displacement map, the zero planes, the FFT normalisation and the crop.
"""
import os

import numpy as np
from scipy.fft import fftn, ifftn

FFT_WORKERS = int(os.environ.get("DDA_FFT_WORKERS", max(1, (os.cpu_count() or 1) - 2)))


def green_element(dx, dy, dz, k):
    """A'(r) for one displacement (float64 3-vector) — the closed form of
    ``mvp_fft.py:78-84``:  e^{ikr}/r [k^2 (rr - I) + (1 - ikr)/r^2 (I - 3 rr)].
    """
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    u = np.array([dx, dy, dz]) / r
    rr = np.outer(u, u)
    eye = np.eye(3)
    return np.exp(1j * k * r) / r * (k ** 2 * (rr - eye) + (1 - 1j * k * r) / r ** 2 * (eye - 3 * rr))


def interaction_tensor(n, lf, k):
    """(2Nx,2Ny,2Nz,3,3) complex128 tensor, ``mvp_fft.py:40-94``.

    Extended index I -> displacement I if I <= N else I - 2N (``:49-51``);
    r = 0 replaced by 1 before dividing (``:62``); origin and every I = N
    plane zeroed (``:87-92``).
    """
    nx, ny, nz = (int(v) for v in n)
    disp = []
    for N in (nx, ny, nz):
        I = np.arange(2 * N)
        disp.append(np.where(I <= N, I, I - 2 * N).astype(np.float64))
    DX, DY, DZ = np.meshgrid(disp[0] * lf, disp[1] * lf, disp[2] * lf, indexing='ij', sparse=True)
    r = np.sqrt(DX ** 2 + DY ** 2 + DZ ** 2)
    rs = np.where(r == 0.0, 1.0, r)
    rhat = np.stack([DX / rs, DY / rs, DZ / rs], axis=-1)
    rr = rhat[..., :, None] * rhat[..., None, :]
    eye = np.eye(3, dtype=np.complex128).reshape(1, 1, 1, 3, 3)
    ph = (np.exp(1j * k * rs) / rs)[..., None, None]
    c2 = ((1.0 - 1j * k * rs) / rs ** 2)[..., None, None]
    A = ph * (k ** 2 * (rr - eye) + c2 * (eye - 3.0 * rr))
    A[0, 0, 0] = 0.0
    A[nx] = 0.0
    A[:, ny] = 0.0
    A[:, :, nz] = 0.0
    return A


def fft_init(n, f, lf, k):
    """FFT of the interaction tensor over the 3 spatial axes (``mvp_fft.py:97-119``)."""
    if f != 3:
        raise ValueError("fft_init: only f=3 (electric dipoles) is supported.")
    return fftn(interaction_tensor(n, lf, k), axes=(0, 1, 2), workers=FFT_WORKERS)


def fft_mvp(n, f, A_tilde, P_block):
    """Off-diagonal block product via zero-padded FFT (``mvp_fft.py:122-173``).

    P_block is (f n,) or (f n, L) on the full cuboid (row = f*cell + c);
    returns the same shape.  Forward FFT unscaled, inverse scaled 1/M.
    """
    nx, ny, nz = (int(v) for v in n)
    one = P_block.ndim == 1
    P = P_block[:, None] if one else P_block
    L = P.shape[1]
    pad = np.zeros((2 * nx, 2 * ny, 2 * nz, f, L), dtype=np.complex128)
    pad[:nx, :ny, :nz] = P.reshape(nx, ny, nz, f, L)
    Pt = fftn(pad, axes=(0, 1, 2), workers=FFT_WORKERS)
    Yt = A_tilde @ Pt
    Y = ifftn(Yt, axes=(0, 1, 2), workers=FFT_WORKERS)
    out = Y[:nx, :ny, :nz].reshape(f * nx * ny * nz, L)
    return out[:, 0] if one else out
