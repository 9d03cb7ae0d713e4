"""bdda_amd — MI355X-native block-DDA solver (HIP kernels behind a C ABI).

The drop-in modules mirroring the reference block-DDA_Py v0.3.0 API live
next to this package (``mvp_fft``, ``bl_krylov``, ``bl_dda``,
``shape_model``); put ``block-dda_py_amd/`` on ``sys.path`` and import them
exactly as with the reference.
"""
from ._lib import BddaError, LIB_PATH, load  # noqa: F401
from .plan import Plan, inside_mask, nearest_distance, require_gpu  # noqa: F401

__version__ = "0.1.0"
