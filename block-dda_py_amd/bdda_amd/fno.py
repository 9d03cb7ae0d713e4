"""Learned warm start: a 3-D Fourier neural operator on the device.

Maps, per (particle, orientation) sample on a G^3 cube, the input channels
(occupancy, per-axis susceptibility chi = eps_r - 1, incident field
E_inc = e0 exp(i k u.r) inside the particle, k * lattice spacing; 14 real
channels) to the local field E_loc = X / alpha (3 complex = 6 real channels).
The block solver then starts from X0 = alpha * E_loc (``bdda_solve``'s X0,
with its fallback to a zero start when the guess is worse than zero).

Architecture: lift 14 -> width, ``layers`` Fourier layers
v <- GELU(W v + b + K v) (no GELU after the last one), projection
width -> hidden (GELU) -> 6.  K v keeps ``modes`` frequencies per axis
(0..modes/2-1 and the negative ones), multiplies them by complex per-mode
(width x width) weights and transforms back.

Every linear stage is a call of libbdda's batched strided bf16-MFMA GEMM
(``bdda_gemm_bf16``): the pointwise layers, the truncated DFTs (products with
fixed DFT matrices along z, y, x), the per-mode weights, and all transposes
of the backward pass.  Complex tensors are pairs of float32 planes.  GELU,
bias, loss and Adam are libbdda kernels too; torch only allocates.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from .plan import require_gpu

F32 = torch.float32


def _p(t, off=0):
    return ctypes.c_void_p(t.data_ptr() + 4 * off)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def mode_indices(G, K):
    """Kept frequencies per axis: 0..K/2-1 and G-K/2..G-1."""
    h = K // 2
    return np.concatenate([np.arange(h), np.arange(G - (K - h), G)])


class Gemm:
    """C = alpha A B + beta C over `batch` products with element strides.

    Operands are (tensor, offset, (batch stride, row stride, col stride)).
    `bias` (M floats) is added to every row of every product, then act=1
    applies GELU, in the product's own epilogue.
    """

    def __init__(self):
        self.lib = _lib.load()

    def __call__(self, batch, M, N, K, A, B, C, alpha=1.0, beta=0.0, bias=None, act=0):
        (ta, oa, sa), (tb, ob, sb), (tc, oc, sc) = A, B, C
        args = (int(batch), int(M), int(N), int(K), ctypes.c_float(alpha), _p(ta, oa), sa[0], sa[1], sa[2],
                _p(tb, ob), sb[0], sb[1], sb[2], ctypes.c_float(beta), _p(tc, oc), sc[0], sc[1], sc[2])
        if bias is None and not act:
            check(self.lib.bdda_gemm_bf16(*args, _stream()), "gemm_bf16")
        else:
            bp = ctypes.c_void_p(None) if bias is None else _p(bias)
            check(self.lib.bdda_gemm_bf16_ex(*args, bp, int(act), _stream()), "gemm_bf16")


class FNO:
    """Fourier neural operator with hand-written forward and backward passes."""

    def __init__(self, G=32, width=32, modes=12, layers=4, cin=14, cout=6, hidden=64, seed=0, params=None):
        require_gpu()
        self.lib = _lib.load()
        self.G, self.C, self.K, self.nl = G, width, modes, layers
        self.cin, self.cout, self.hid = cin, cout, hidden
        self.S = G ** 3
        self.Km = modes ** 3
        self.gemm = Gemm()
        ks = mode_indices(G, modes)
        n = np.arange(G)
        F = np.exp(-2j * np.pi * np.outer(ks, n) / G)           # (K, G)
        Fi = np.exp(2j * np.pi * np.outer(n, ks) / G) / G        # (G, K)
        dev = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=F32, device="cuda")
        self.Fr, self.Fim = dev(F.real), dev(F.imag)
        self.Fir, self.Fii = dev(Fi.real), dev(Fi.imag)
        # the z transforms of the forward pass as one product each way, so the
        # activations are read (and the output written) once: F's real and
        # imaginary rows stacked, (2K, G); Fi's real and negated imaginary
        # columns side by side, (G, 2K)
        self.Fz = dev(np.concatenate([F.real, F.imag], axis=0))
        self.Fzi = dev(np.concatenate([Fi.real, -Fi.imag], axis=1))
        self.params = {}
        if params is None:
            rng = np.random.default_rng(seed)

            def lin(co, ci):
                bnd = 1.0 / np.sqrt(ci)
                return rng.uniform(-bnd, bnd, (co, ci)), rng.uniform(-bnd, bnd, co)

            w, b = lin(width, cin)
            self.params["lift.W"], self.params["lift.b"] = w, b
            sc = 1.0 / (width * width)
            for l in range(layers):
                w, b = lin(width, width)
                self.params[f"l{l}.W"], self.params[f"l{l}.b"] = w, b
                self.params[f"l{l}.Rr"] = sc * rng.uniform(0, 1, (self.Km, width, width))
                self.params[f"l{l}.Ri"] = sc * rng.uniform(0, 1, (self.Km, width, width))
            w, b = lin(hidden, width)
            self.params["p1.W"], self.params["p1.b"] = w, b
            w, b = lin(cout, hidden)
            self.params["p2.W"], self.params["p2.b"] = w, b
            params = self.params
        self.params = {k: dev(np.asarray(v, dtype=np.float32)) for k, v in params.items()}
        self.adam = None
        self.step_count = 0

    # ------------------------------------------------------------ helpers
    def _empty(self, *shape):
        return torch.empty(shape, dtype=F32, device="cuda")

    def _pointwise(self, v, W, out, B, ci, co, beta=0.0, bias=None, act=0):
        """out[b] (co, S) = W (co, ci) @ v[b] (ci, S)  (+ beta out) (+ bias),
        then GELU if act."""
        S = self.S
        self.gemm(B, co, S, ci, (W, 0, (0, ci, 1)), (v, 0, (ci * S, S, 1)), (out, 0, (co * S, S, 1)), beta=beta,
                  bias=bias, act=act)

    def _pointwise_bwd(self, v, W, dout, dv, dW, B, ci, co, dv_beta):
        S = self.S
        if dv is not None:
            self.gemm(B, ci, S, co, (W, 0, (0, 1, ci)), (dout, 0, (co * S, S, 1)), (dv, 0, (ci * S, S, 1)),
                      beta=dv_beta)
        # dW = sum_b dout[b] v[b]^T, split over chunks of S then summed by a ones-GEMM
        nch = max(1, S // 4096)
        while S % nch:
            nch -= 1
        Sc = S // nch
        part = self._empty(B, nch, co, ci)
        for b in range(B):
            self.gemm(nch, co, ci, Sc, (dout, b * co * S, (Sc, S, 1)), (v, b * ci * S, (Sc, 1, S)),
                      (part, b * nch * co * ci, (co * ci, ci, 1)))
        ones = torch.ones(B * nch, dtype=F32, device="cuda")
        self.gemm(1, 1, co * ci, B * nch, (ones, 0, (0, 0, 1)), (part, 0, (0, co * ci, 1)), (dW, 0, (0, 0, 1)))

    def _bias_add(self, y, b, B, C):
        check(self.lib.bdda_bias_add(_p(y), _p(b), B, C, self.S, _stream()), "bias_add")

    def _bias_grad(self, dy, db, B, C):
        check(self.lib.bdda_bias_grad(_p(dy), _p(db), B, C, self.S, _stream()), "bias_grad")

    def _gelu(self, x, y):
        check(self.lib.bdda_gelu(_p(x), _p(y), x.numel(), _stream()), "gelu")

    def _gelu_bwd(self, x, dy, dx):
        check(self.lib.bdda_gelu_bwd(_p(x), _p(dy), _p(dx), x.numel(), _stream()), "gelu_bwd")

    # ----------------------------------------------------- spectral layer
    def _cmm(self, batch, M, N, K, Ar, Ai, sa, Br, Bi, sb, Cr, Ci, sc, conjA=False):
        """Complex (Ar + i Ai)(Br + i Bi) (conjA: conj(A)) into (Cr, Ci).
        Each part is a tensor or a (tensor, element offset) pair."""
        g = self.gemm
        s = -1.0 if conjA else 1.0
        Ar, Ai, Br, Bi, Cr, Ci = (x if isinstance(x, tuple) else (x, 0) for x in (Ar, Ai, Br, Bi, Cr, Ci))
        g(batch, M, N, K, (*Ar, sa), (*Br, sb), (*Cr, sc))
        g(batch, M, N, K, (*Ai, sa), (*Bi, sb), (*Cr, sc), alpha=-s, beta=1.0)
        g(batch, M, N, K, (*Ar, sa), (*Bi, sb), (*Ci, sc))
        g(batch, M, N, K, (*Ai, sa), (*Br, sb), (*Ci, sc), alpha=s, beta=1.0)

    def _spectral_fwd(self, v, Rr, Ri, B, out):
        G, K, C, Km = self.G, self.K, self.C, self.Km
        g = self.gemm
        # z: real (G; B C G, G) x [F_re; F_im]^T -> Z (G, B C G, 2K), y outermost,
        # each row its K real parts, then its K imaginary parts
        Q = B * C * G
        Xz = self._empty(G * Q * 2 * K)
        g(G, Q, 2 * K, G, (v, 0, (G, G * G, 1)), (self.Fz, 0, (0, 1, G)), (Xz, 0, (Q * 2 * K, 2 * K, 1)))
        # y: one wide product [F_re; F_im] (2K, G) x Z (G, Q 2K), then the
        # complex combination into (Q, K, K) planes
        R = self._empty(2 * K * Q * 2 * K)
        g(1, 2 * K, Q * 2 * K, G, (self.Fz, 0, (0, G, 1)), (Xz, 0, (0, Q * 2 * K, 1)), (R, 0, (0, Q * 2 * K, 1)))
        Xyr, Xyi = self._empty(Q * K * K), self._empty(Q * K * K)
        check(self.lib.bdda_fno_ycombine(_p(R), _p(Xyr), _p(Xyi), Q, K, _stream()), "fno_ycombine")
        # x
        Xhr, Xhi = self._empty(B * C * Km), self._empty(B * C * Km)
        self._cmm(B * C, K, K * K, G, self.Fr, self.Fim, (0, G, 1), Xyr, Xyi, (G * K * K, K * K, 1), Xhr, Xhi,
                  (Km, K * K, 1))
        # per-mode weights: Y[b][co] = sum_ci X[b][ci] R[m][ci][co]
        Yhr, Yhi = self._empty(B * C * Km), self._empty(B * C * Km)
        self._cmm(Km, B, C, C, Xhr, Xhi, (1, C * Km, Km), Rr, Ri, (C * C, C, 1), Yhr, Yhi, (1, C * Km, Km))
        # inverse x
        Yxr, Yxi = self._empty(B * C * G * K * K), self._empty(B * C * G * K * K)
        self._cmm(B * C, G, K * K, K, self.Fir, self.Fii, (0, K, 1), Yhr, Yhi, (Km, K * K, 1), Yxr, Yxi,
                  (G * K * K, K * K, 1))
        # inverse y, into rows of K real parts then K imaginary parts
        Yy = self._empty(B * C * G * G * 2 * K)
        self._cmm(B * C * G, G, K, K, self.Fir, self.Fii, (0, K, 1), Yxr, Yxi, (K * K, K, 1), (Yy, 0), (Yy, K),
                  (G * 2 * K, 2 * K, 1))
        # inverse z, real part: (B C G G, 2K) x [Fi_re, -Fi_im]^T
        g(1, B * C * G * G, G, 2 * K, (Yy, 0, (0, 2 * K, 1)), (self.Fzi, 0, (0, 1, 2 * K)), (out, 0, (0, G, 1)))
        return Xhr, Xhi

    def _spectral_bwd(self, dout, Xhr, Xhi, Rr, Ri, B, dv, dRr, dRi):
        """dv += (adjoint of the spectral layer) dout; dR = per-mode weight gradient."""
        G, K, C, Km = self.G, self.K, self.C, self.Km
        g = self.gemm
        dYyr, dYyi = self._empty(B * C * G * G * K), self._empty(B * C * G * G * K)
        g(1, B * C * G * G, K, G, (dout, 0, (0, G, 1)), (self.Fir, 0, (0, K, 1)), (dYyr, 0, (0, K, 1)))
        g(1, B * C * G * G, K, G, (dout, 0, (0, G, 1)), (self.Fii, 0, (0, K, 1)), (dYyi, 0, (0, K, 1)), alpha=-1.0)
        # adjoint of inverse y: Fi^H along y
        dYxr, dYxi = self._empty(B * C * G * K * K), self._empty(B * C * G * K * K)
        self._cmm(B * C * G, K, K, G, self.Fir, self.Fii, (0, 1, K), dYyr, dYyi, (G * K, K, 1), dYxr, dYxi,
                  (K * K, K, 1), conjA=True)
        # adjoint of inverse x
        dYhr, dYhi = self._empty(B * C * Km), self._empty(B * C * Km)
        self._cmm(B * C, K, K * K, G, self.Fir, self.Fii, (0, 1, K), dYxr, dYxi, (G * K * K, K * K, 1), dYhr, dYhi,
                  (Km, K * K, 1), conjA=True)
        # weight gradient dR[m] = X^H dY  (M = ci, N = co, K = b)
        self._cmm(Km, C, C, B, Xhr, Xhi, (1, Km, C * Km), dYhr, dYhi, (1, C * Km, Km), dRr, dRi, (C * C, C, 1),
                  conjA=True)
        # dX = dY R^H  (per mode, B x C): conj(R)^T
        dXhr, dXhi = self._empty(B * C * Km), self._empty(B * C * Km)
        g(Km, B, C, C, (dYhr, 0, (1, C * Km, Km)), (Rr, 0, (C * C, 1, C)), (dXhr, 0, (1, C * Km, Km)))
        g(Km, B, C, C, (dYhi, 0, (1, C * Km, Km)), (Ri, 0, (C * C, 1, C)), (dXhr, 0, (1, C * Km, Km)), beta=1.0)
        g(Km, B, C, C, (dYhi, 0, (1, C * Km, Km)), (Rr, 0, (C * C, 1, C)), (dXhi, 0, (1, C * Km, Km)))
        g(Km, B, C, C, (dYhr, 0, (1, C * Km, Km)), (Ri, 0, (C * C, 1, C)), (dXhi, 0, (1, C * Km, Km)), alpha=-1.0,
          beta=1.0)
        # adjoint of forward x: F^H
        dXyr, dXyi = self._empty(B * C * G * K * K), self._empty(B * C * G * K * K)
        self._cmm(B * C, G, K * K, K, self.Fr, self.Fim, (0, 1, G), dXhr, dXhi, (Km, K * K, 1), dXyr, dXyi,
                  (G * K * K, K * K, 1), conjA=True)
        # adjoint of forward y
        dXzr, dXzi = self._empty(B * C * G * G * K), self._empty(B * C * G * G * K)
        self._cmm(B * C * G, G, K, K, self.Fr, self.Fim, (0, 1, G), dXyr, dXyi, (K * K, K, 1), dXzr, dXzi,
                  (G * K, K, 1), conjA=True)
        # adjoint of the real -> complex z transform: dv += dXz_re F_re + dXz_im F_im
        g(1, B * C * G * G, G, K, (dXzr, 0, (0, K, 1)), (self.Fr, 0, (0, G, 1)), (dv, 0, (0, G, 1)), beta=1.0)
        g(1, B * C * G * G, G, K, (dXzi, 0, (0, K, 1)), (self.Fim, 0, (0, G, 1)), (dv, 0, (0, G, 1)), beta=1.0)

    # ------------------------------------------------------- forward/backward
    def forward(self, x, keep=False):
        """x: (B, cin, S) float32 device tensor -> (B, cout, S)."""
        P = self.params
        B = x.shape[0]
        C, S = self.C, self.S
        cache = {"x": x}
        # biases (and, where no pre-activation is kept for the backward
        # pass, the GELUs) run in the pointwise products' epilogues
        v = self._empty(B, C, S)
        self._pointwise(x, P["lift.W"], v, B, self.cin, C, bias=P["lift.b"])
        for l in range(self.nl):
            last = l == self.nl - 1
            z = self._empty(B, C, S)
            Xh = self._spectral_fwd(v, P[f"l{l}.Rr"], P[f"l{l}.Ri"], B, z)
            self._pointwise(v, P[f"l{l}.W"], z, B, C, C, beta=1.0, bias=P[f"l{l}.b"], act=int(not keep and not last))
            if keep:
                cache[f"v{l}"], cache[f"z{l}"], cache[f"Xh{l}"] = v, z, Xh
            if keep and not last:
                vn = self._empty(B, C, S)
                self._gelu(z, vn)
                v = vn
            else:
                v = z
        h = self._empty(B, self.hid, S)
        self._pointwise(v, P["p1.W"], h, B, C, self.hid, bias=P["p1.b"], act=int(not keep))
        if keep:
            hg = self._empty(B, self.hid, S)
            self._gelu(h, hg)
        else:
            hg = h
        y = self._empty(B, self.cout, S)
        self._pointwise(hg, P["p2.W"], y, B, self.hid, self.cout, bias=P["p2.b"])
        if keep:
            cache.update(vL=v, h=h, hg=hg)
            self.cache = cache
        return y

    def backward(self, dy):
        """Gradients of every parameter for the upstream gradient dy (B, cout, S)."""
        P, c = self.params, self.cache
        B = dy.shape[0]
        C, S = self.C, self.S
        gr = {k: torch.empty_like(v) for k, v in P.items()}
        self._bias_grad(dy, gr["p2.b"], B, self.cout)
        dhg = self._empty(B, self.hid, S)
        self._pointwise_bwd(c["hg"], P["p2.W"], dy, dhg, gr["p2.W"], B, self.hid, self.cout, 0.0)
        dh = self._empty(B, self.hid, S)
        self._gelu_bwd(c["h"], dhg, dh)
        self._bias_grad(dh, gr["p1.b"], B, self.hid)
        dv = self._empty(B, C, S)
        self._pointwise_bwd(c["vL"], P["p1.W"], dh, dv, gr["p1.W"], B, C, self.hid, 0.0)
        for l in reversed(range(self.nl)):
            if l < self.nl - 1:
                dz = self._empty(B, C, S)
                self._gelu_bwd(c[f"z{l}"], dv, dz)
            else:
                dz = dv
            self._bias_grad(dz, gr[f"l{l}.b"], B, C)
            dvp = self._empty(B, C, S)
            self._pointwise_bwd(c[f"v{l}"], P[f"l{l}.W"], dz, dvp, gr[f"l{l}.W"], B, C, C, 0.0)
            Xhr, Xhi = c[f"Xh{l}"]
            self._spectral_bwd(dz, Xhr, Xhi, P[f"l{l}.Rr"], P[f"l{l}.Ri"], B, dvp, gr[f"l{l}.Rr"], gr[f"l{l}.Ri"])
            dv = dvp
        self._bias_grad(dv, gr["lift.b"], B, C)
        self._pointwise_bwd(c["x"], P["lift.W"], dv, None, gr["lift.W"], B, self.cin, C, 0.0)
        return gr

    def adam_step(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        if self.adam is None:
            self.adam = {k: (torch.zeros_like(v), torch.zeros_like(v)) for k, v in self.params.items()}
        self.step_count += 1
        for k, p in self.params.items():
            m, v = self.adam[k]
            check(self.lib.bdda_adam(_p(p), _p(grads[k]), _p(m), _p(v), p.numel(), ctypes.c_float(lr),
                                     ctypes.c_float(b1), ctypes.c_float(b2), ctypes.c_float(eps), self.step_count,
                                     _stream()), "adam")

    # ----------------------------------------------------------------- io
    def state(self):
        return {k: v.cpu().numpy() for k, v in self.params.items()}

    def save(self, path):
        meta = np.array([self.G, self.C, self.K, self.nl, self.cin, self.cout, self.hid], dtype=np.int64)
        np.savez_compressed(path, __meta__=meta, **{k: v.astype(np.float16) for k, v in self.state().items()})

    def at_grid(self, G):
        """The same network on a G^3 cube: the spectral weights are indexed by
        mode (0..5 and G-6..G-1 per axis), so they carry over to any G >= 2
        modes; only the truncated DFT matrices change."""
        return FNO(G=G, width=self.C, modes=self.K, layers=self.nl, cin=self.cin, cout=self.cout,
                   hidden=self.hid, params={k: np.array(v, dtype=np.float32) for k, v in self.state().items()})

    @classmethod
    def load(cls, path):
        d = np.load(path)
        G, C, K, nl, cin, cout, hid = (int(v) for v in d["__meta__"])
        params = {k: d[k].astype(np.float32) for k in d.files if k != "__meta__"}
        return cls(G=G, width=C, modes=K, layers=nl, cin=cin, cout=cout, hidden=hid, params=params)
