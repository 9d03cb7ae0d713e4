"""CPU oracle for the block-DDA solve path — TEST INFRASTRUCTURE ONLY.

This package is a float64 NumPy/SciPy restatement of the reference
block-DDA_Py v0.3.0 algorithms on the solve path (GRE voxelisation, Goodman
FFT interaction operator, block-COCG-rq / block-BiCGSTAB, DDA observables,
Mie theory).  Every function cites the reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker / CPU baseline.  The
product (``block-dda_py_amd/``) never imports, links or executes anything in
this package.

Parity status of each module (how each function is pinned) is listed in the
module docstrings; a function without such a check says "parity unpinned".
"""
