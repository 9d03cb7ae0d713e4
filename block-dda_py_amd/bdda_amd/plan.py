"""Device plan: one lattice, its Green tensor spectrum, its occupied target.

Thin object layer over the C ABI.  Device buffers are torch tensors on the
current ROCm device (torch is used for allocation and streams only); every
computation runs in libbdda's HIP kernels.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr


def require_gpu():
    if not torch.cuda.is_available():
        raise _lib.BddaError("bdda_amd needs a ROCm GPU (MI355X); there is no CPU fallback")


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def to_device(a, dtype=torch.complex128):
    """numpy array or tensor -> contiguous device tensor of the given dtype."""
    if isinstance(a, torch.Tensor):
        t = a.to(device="cuda", dtype=dtype)
    else:
        t = torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)
    return t.contiguous()


class Plan:
    """A libbdda plan over an (nx, ny, nz) cuboid lattice."""

    def __init__(self, n):
        require_gpu()
        self.lib = _lib.load()
        self.n = tuple(int(v) for v in n)
        h = ctypes.c_void_p()
        check(self.lib.bdda_plan_create(self.n[0], self.n[1], self.n[2], ctypes.byref(h)), "plan_create")
        self.h = h
        m = (ctypes.c_int * 3)()
        check(self.lib.bdda_plan_fft_dims(self.h, m), "plan_fft_dims")
        self.m = tuple(m)
        self.ncell = int(np.prod(self.n))
        self.N = None
        self.target_key = None
        self.lf = self.k = None

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.lib.bdda_plan_destroy(h)
            except Exception:
                pass
            self.h = None

    # ------------------------------------------------------------ operator
    def green_init(self, lf, k):
        check(self.lib.bdda_green_init(self.h, float(lf), float(k)), "green_init")
        self.lf, self.k = float(lf), float(k)

    def share_green(self, other):
        """Compute with `other`'s Green spectrum (same cuboid) instead of
        building one."""
        check(self.lib.bdda_green_share(self.h, other.h), "green_share")
        self.lf, self.k = other.lf, other.k

    def green_spectrum(self):
        out = np.empty(self.m + (3, 3), dtype=np.complex128)
        check(self.lib.bdda_green_download(self.h, ptr(out)), "green_download")
        return out

    def column_tile(self, L):
        """Orientations per line-FFT tile that the operator uses for L columns."""
        lb = self.lib.bdda_plan_column_tile(self.h, int(L))
        if lb < 1:
            raise ValueError("column_tile: bad L")
        return lb

    def set_column_tile(self, lb):
        """Fix the column tile (0: automatic)."""
        check(self.lib.bdda_plan_set_column_tile(self.h, int(lb)), "set_column_tile")

    def set_x_kernel(self, kernel):
        """Fix the x pass kernel: 0 the process-wide rule, 1 split-spectrum
        where eligible, 2 one-pass, 3 split-spectrum where balanced."""
        check(self.lib.bdda_plan_set_x_kernel(self.h, int(kernel)), "set_x_kernel")

    def x_kernel(self):
        k = self.lib.bdda_plan_x_kernel(self.h)
        if k < 0:
            raise ValueError("x_kernel: bad plan")
        return k

    def autotune(self, L, precision="fp64"):
        """Time the operator at the candidate column tiles, each with the
        split-spectrum and the one-pass x pass, for L columns and keep the
        fastest pair; returns (tile, x kernel)."""
        lb = ctypes.c_int(0)
        check(self.lib.bdda_plan_autotune(self.h, int(L), _lib.PRECISION[precision], ctypes.byref(lb), _stream()),
              "autotune")
        return lb.value, self.x_kernel()

    def green_tensor(self, lf, k):
        out = np.empty(self.m + (3, 3), dtype=np.complex128)
        check(self.lib.bdda_green_tensor(self.h, float(lf), float(k), ptr(out)), "green_tensor")
        return out

    def mvp_offdiag(self, P, precision="fp64"):
        """Y = A_off P on the full cuboid; P (3n, L) complex128 device tensor."""
        P = to_device(P)
        if P.ndim != 2 or P.shape[0] != 3 * self.ncell:
            raise ValueError(f"P must be (3n, L) with 3n = {3 * self.ncell}")
        Y = torch.empty_like(P)
        check(self.lib.bdda_mvp_offdiag(self.h, ptr(P), ptr(Y), P.shape[1], _lib.PRECISION[precision],
                                        _stream()), "mvp_offdiag")
        return Y

    def profile_enable(self, on=True):
        check(self.lib.bdda_profile_enable(self.h, int(bool(on))), "profile_enable")

    def profile_read(self):
        """{(precision, pass): (ms_sum, count)} since the last read."""
        ms = (ctypes.c_double * 10)()
        cnt = (ctypes.c_int64 * 10)()
        check(self.lib.bdda_profile_read(self.h, ms, cnt), "profile_read")
        names = ["z_fwd", "y_fwd", "x_conv", "y_inv", "z_inv"]
        out = {}
        for i in range(10):
            if cnt[i]:
                out[("fp64" if i < 5 else "fp32", names[i % 5])] = (ms[i], int(cnt[i]))
        return out

    # -------------------------------------------------------------- target
    def set_target(self, address, diag, jpre, pos=None, key=None):
        address = np.ascontiguousarray(address, dtype=np.int64)
        diag = np.ascontiguousarray(diag, dtype=np.complex128)
        jpre = np.ascontiguousarray(jpre, dtype=np.complex128)
        N = address.size
        if diag.size != 3 * N or jpre.size != 3 * N:
            raise ValueError("diag and jpre must have 3N entries")
        posa = None if pos is None else np.ascontiguousarray(pos, dtype=np.float64)
        check(self.lib.bdda_set_target(self.h, ptr(address), N, ptr(posa), ptr(diag), ptr(jpre)), "set_target")
        self.N = N
        # occupied extent (ix, iy, iz bounds of the occupied cells): the
        # operator's passes skip the empty lines and rows beyond it
        ny, nz = self.n[1], self.n[2]
        self.occupied_extent = (int(address.max() // (ny * nz)) + 1, int((address // nz % ny).max()) + 1,
                                int((address % nz).max()) + 1)
        # z lines (ix, iy) holding an occupied cell: the z passes run only these
        self.occupied_lines = int(np.unique(address // nz).size)
        self.target_key = key

    def block_mvp(self, S):
        S = to_device(S)
        AS = torch.empty_like(S)
        check(self.lib.bdda_block_mvp(self.h, ptr(S), ptr(AS), S.shape[1], _stream()), "block_mvp")
        return AS

    def solve(self, B, solver="bicgstab", precision="fp64", tol=1e-2, itermax=25, X0=None, refine_every=20,
              true_residual=True):
        """Returns (X, iter_fin, err_fin, true_res, history, warm_start_used);
        self.last_solve_info then holds the mixed-precision refinement
        cycles and whether the solve fell back to complex128.
        A float64 solve stops on its recursive residual, as the reference's
        solvers do; true_residual=False skips the extra operator application
        that checks it (true_res is then None).  A mixed solve's err_fin is
        its true residual already."""
        B = to_device(B)
        L = B.shape[1]
        X = torch.empty_like(B)
        X0d = None if X0 is None else to_device(X0)
        it = ctypes.c_int(0)
        err = ctypes.c_double(0.0)
        tr = ctypes.c_double(0.0)
        flags = ctypes.c_int(0)
        hist = np.zeros(max(int(itermax), 1), dtype=np.float64)
        check(self.lib.bdda_solve(self.h, _lib.SOLVER[solver], _lib.PRECISION[precision], ptr(B), ptr(X0d),
                                  ptr(X), L, float(tol), int(itermax), int(refine_every), ctypes.byref(it),
                                  ctypes.byref(err), ctypes.byref(tr) if true_residual else None,
                                  hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(flags),
                                  _stream()), "solve")
        nh = flags.value >> 8
        self.last_solve_info = {"refine_cycles": (flags.value >> 2) & 63,
                                "fallback_complex128": bool(flags.value & 2)}
        trv = tr.value if true_residual else (err.value if precision == "mixed" else None)
        return X, it.value, err.value, trv, hist[:nh].tolist(), bool(flags.value & 1)

    def true_residual(self, B, X):
        """||jpre (B - A X)||_F / ||jpre B||_F with one complex128 operator
        application (B, X: (3N, L) complex128 device tensors)."""
        B, X = to_device(B), to_device(X)
        r = ctypes.c_double(0.0)
        check(self.lib.bdda_true_residual(self.h, ptr(B), ptr(X), B.shape[1], ctypes.byref(r), _stream()),
              "true_residual")
        return r.value

    def project_guess(self, B, X0):
        """X0 (3N, L) complex128 device tensor, in place -> X0 C minimising
        ||jpre (B - A X0 C)||_F; returns the relative residuals (before, after)."""
        B = to_device(B)
        if not (isinstance(X0, torch.Tensor) and X0.is_cuda and X0.dtype == torch.complex128
                and X0.is_contiguous() and tuple(X0.shape) == tuple(B.shape)):
            raise ValueError("X0 must be a contiguous complex128 device tensor shaped like B")
        before = ctypes.c_double(0.0)
        after = ctypes.c_double(0.0)
        check(self.lib.bdda_project_guess(self.h, ptr(B), ptr(X0), B.shape[1], ctypes.byref(before),
                                          ctypes.byref(after), _stream()), "project_guess")
        return before.value, after.value

    def rhs(self, u, e0, k):
        u = np.ascontiguousarray(u, dtype=np.float64)
        e0 = np.ascontiguousarray(e0, dtype=np.complex128)
        L = u.shape[0]
        B = torch.empty((3 * self.N, L), dtype=torch.complex128, device="cuda")
        check(self.lib.bdda_rhs(self.h, ptr(u), ptr(e0), float(k), L, ptr(B), _stream()), "rhs")
        return B

    def observables(self, X, geo, epsfac, k):
        X = to_device(X)
        geo = np.ascontiguousarray(geo, dtype=np.float64)
        epsfac = np.ascontiguousarray(epsfac, dtype=np.complex128)
        L = X.shape[1]
        out = np.empty((L, 5), dtype=np.complex128)
        check(self.lib.bdda_observables(self.h, ptr(X), L, ptr(geo), ptr(epsfac), float(k), ptr(out), _stream()),
              "observables")
        return out


def nearest_distance(grid_points, surf_points):
    lib = _lib.load()
    require_gpu()
    g = np.ascontiguousarray(grid_points, dtype=np.float64)
    s = np.ascontiguousarray(surf_points, dtype=np.float64)
    out = np.empty(g.shape[0], dtype=np.float64)
    check(lib.bdda_nearest_distance(ptr(g), g.shape[0], ptr(s), s.shape[0], ptr(out)), "nearest_distance")
    return out


def inside_mask(dist, lattice_n, thresh):
    lib = _lib.load()
    require_gpu()
    d = np.ascontiguousarray(dist, dtype=np.float64)
    nx, ny, nz = (int(v) for v in lattice_n)
    out = np.empty(nx * ny * nz, dtype=np.uint8)
    check(lib.bdda_inside_mask(ptr(d), nx, ny, nz, float(thresh), ptr(out)), "inside_mask")
    return out.astype(bool)
