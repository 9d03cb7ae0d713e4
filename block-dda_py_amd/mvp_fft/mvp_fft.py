"""Goodman doubled-grid FFT interaction operator on the GPU.

Drop-in for the reference ``mvp_fft/mvp_fft.py`` (v0.3.0): same function
names, arguments, shapes and errors.  ``fft_init`` returns a device-resident
``GreenTensor`` handle instead of a host array; ``fft_mvp`` accepts it, and
``numpy.asarray(handle)`` gives the reference's (2Nx, 2Ny, 2Nz, 3, 3) array
when every doubled axis length 2N is 5-smooth (the device then uses exactly
the reference grid).  All arithmetic runs in libbdda's HIP kernels.
"""
import os
from collections import OrderedDict

import numpy as np
import torch

from bdda_amd.plan import Plan, to_device

# Kept for interface compatibility with mvp_fft.py:9-10; the GPU path has no
# host FFT threads.
_FFT_WORKERS = int(os.environ.get("DDA_FFT_WORKERS", max(1, (os.cpu_count() or 1) - 2)))


class GreenTensor:
    """Device-resident FFT of the interaction tensor (what ``fft_init`` returns).

    Holds the libbdda plan: complex128 and complex64 half spectra of the 6
    unique components on the device, shared with every other handle of the
    same cuboid, lattice spacing and wavenumber while it is cached.
    ``shape`` is the convolution grid (m0, m1, m2, 3, 3); m = 2N when 2N is
    5-smooth, else the smallest 5-smooth length >= 2N - 1 (an exact linear
    convolution either way).
    """

    dtype = np.dtype(np.complex128)

    # Spectra built so far, by (cuboid, lattice_lf, k): a few spectrum-only
    # plans kept alive so that the next instance with the same lattice and
    # wavenumber (another m_p, another set of orientations) shares the
    # spectrum instead of rebuilding it.  A_off does not depend on m_p.
    cache_size = 4
    _cache = OrderedDict()
    builds = 0

    def __init__(self, n, lf, k, plan=None):
        self.n = tuple(int(v) for v in n)
        self.lf, self.k = float(lf), float(k)
        self.plan = plan if plan is not None else Plan(self.n)
        key = (self.n, self.lf, self.k)
        holder = GreenTensor._cache.get(key)
        if holder is None:
            holder = Plan(self.n)
            holder.green_init(self.lf, self.k)
            GreenTensor.builds += 1
            if GreenTensor.cache_size > 0:
                GreenTensor._cache[key] = holder
                while len(GreenTensor._cache) > GreenTensor.cache_size:
                    GreenTensor._cache.popitem(last=False)
        else:
            GreenTensor._cache.move_to_end(key)
        self.plan.share_green(holder)

    @classmethod
    def clear_cache(cls):
        cls._cache.clear()

    @property
    def shape(self):
        return tuple(self.plan.m) + (3, 3)

    @property
    def ndim(self):
        return 5

    def __array__(self, dtype=None, copy=None):
        if tuple(self.plan.m) != tuple(2 * v for v in self.n):
            raise ValueError("the device grid is %s, not the reference's 2N grid %s; "
                             "the spectrum on the 2N grid is not available" %
                             (self.plan.m, tuple(2 * v for v in self.n)))
        a = self.plan.green_spectrum()
        return a if dtype is None else a.astype(dtype)


def build_interaction_tensor(n, lf, k):
    """Real-space A'(r) on the doubled grid (mvp_fft.py:13-94), built on the
    device.  Returns a (2Nx, 2Ny, 2Nz, 3, 3) complex128 array when every 2N is
    5-smooth (otherwise the array is on the device's convolution grid)."""
    return Plan(n).green_tensor(lf, k)


def fft_init(n, f, lf, k):
    """Device FFT of the interaction tensor (mvp_fft.py:97-119)."""
    if f != 3:
        raise ValueError("fft_init: only f=3 (electric dipoles) is supported.")
    return GreenTensor(n, lf, k)


def fft_mvp(n, f, A_tensor_tilde, P_block, precision="fp64"):
    """Off-diagonal product A_off @ P_block on the full cuboid (mvp_fft.py:122-173).

    P_block: (3n,) or (3n, L), numpy or torch.  Returns the same shape and
    kind (numpy in -> numpy out; a torch tensor stays on the device).
    ``precision='mixed'`` runs the transforms in complex64.
    """
    if f != 3:
        raise ValueError("fft_mvp: only f=3 (electric dipoles) is supported.")
    if not isinstance(A_tensor_tilde, GreenTensor):
        raise TypeError("A_tensor_tilde must be the handle returned by this module's fft_init")
    if tuple(int(v) for v in n) != A_tensor_tilde.n:
        raise ValueError("lattice n does not match the one fft_init was called with")
    is_torch = isinstance(P_block, torch.Tensor)
    one = P_block.ndim == 1
    P = P_block[:, None] if one else P_block
    Y = A_tensor_tilde.plan.mvp_offdiag(to_device(P), precision)
    if one:
        Y = Y[:, 0]
    return Y if is_torch else Y.cpu().numpy()
