"""Float64 CPU Fourier neural operator — oracle for the learned warm start.

The reference has no learned component; this module is the CPU float64
statement of the network that ``block-dda_py_amd/bdda_amd/fno.py`` runs on
the GPU with bf16 MFMA GEMMs.  It is written independently of the GPU
formulation: spectral layers use ``numpy.fft.fftn`` on the full grid, keep
the modes 0..K/2-1 and -(K-K/2)..-1 per axis, multiply by the per-mode
complex weights and transform back with ``ifftn`` (real part); the GPU
version uses truncated DFT matrices along z, y, x instead.

Pinned by: ``numpy.fft`` (the spectral layer equals a direct DFT sum on a
tiny grid, tested), and finite differences of ``loss`` for the gradient
checks of the GPU backward pass.
"""
import numpy as np
from scipy.special import erf


def mode_indices(G, K):
    h = K // 2
    return np.concatenate([np.arange(h), np.arange(G - (K - h), G)])


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def pointwise(W, b, v):
    return np.einsum('oc,bcs->bos', W, v) + b[None, :, None]


def spectral(v, Rr, Ri, G, K):
    """v: (B, C, G^3) real; R: (K^3, C, C) -> (B, C, G^3) real."""
    B, C, _ = v.shape
    ks = mode_indices(G, K)
    Vh = np.fft.fftn(v.reshape(B, C, G, G, G), axes=(2, 3, 4))
    Vh = Vh[:, :, ks][:, :, :, ks][:, :, :, :, ks]
    R = (Rr + 1j * Ri).reshape(K, K, K, C, C)
    Yh = np.einsum('bcxyz,xyzcd->bdxyz', Vh, R)
    full = np.zeros((B, C, G, G, G), dtype=np.complex128)
    full[:, :, ks[:, None, None], ks[None, :, None], ks[None, None, :]] = Yh
    return np.fft.ifftn(full, axes=(2, 3, 4)).real.reshape(B, C, G ** 3)


def forward(p, x, G, K, layers):
    """p: dict of float64 arrays named as in the GPU model; x (B, cin, G^3)."""
    v = pointwise(p["lift.W"], p["lift.b"], x)
    for l in range(layers):
        z = spectral(v, p[f"l{l}.Rr"], p[f"l{l}.Ri"], G, K) + pointwise(p[f"l{l}.W"], p[f"l{l}.b"], v)
        v = gelu(z) if l < layers - 1 else z
    h = gelu(pointwise(p["p1.W"], p["p1.b"], v))
    return pointwise(p["p2.W"], p["p2.b"], h)


def loss(p, x, w, G, K, layers):
    """Linear functional sum(w * forward) (its gradient is the backward pass
    of upstream gradient w)."""
    return float(np.sum(w * forward(p, x, G, K, layers)))
