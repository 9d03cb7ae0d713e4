"""Orientation draws — oracle.

* ``euler_run_dda``  <- ``run_dda._run_dda`` (``run_dda.py:46-54``): one
  ``default_rng(seed)`` stream, alpha (L) then beta (L) then gamma (L), uniform
  in [0,2pi), [0,pi), [0,2pi) radians.  Each retry draws the next 3L values.
* ``euler_test_dda`` <- ``test_dda.ipynb`` cell 5: degrees from
  ``default_rng(12345)`` in the same order, then ``np.radians``.

Pinned by: the Euler angles stored in the reference's HDF5 fixture (first
try, seed 12346) and the angles printed in ``test_dda.ipynb``.
"""
import numpy as np


def euler_run_dda(num, seed=12346, tries=1):
    rng = np.random.default_rng(seed)
    out = None
    for _ in range(tries):
        out = np.column_stack([rng.uniform(0, 2 * np.pi, num),
                               rng.uniform(0, np.pi, num),
                               rng.uniform(0, 2 * np.pi, num)])
    return out


def euler_test_dda(num=10, seed=12345):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 360, num).reshape(num, 1)
    b = rng.uniform(0, 180, num).reshape(num, 1)
    g = rng.uniform(0, 360, num).reshape(num, 1)
    return np.radians(np.hstack((a, b, g)))
