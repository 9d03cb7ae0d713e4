// Small dense complex float64 linear algebra on the host for the L x L
// recurrence coefficients of the block-Krylov solvers (declarations and
// conventions in host_la.h).  Plain C++ (built with the host compiler), so
// each routine can carry an AVX2/FMA clone that the loader selects at
// run time, with a generic x86-64 fallback.
#include "host_la.h"

#include <cmath>
#include <utility>

namespace bdda {
namespace la {
namespace {

// complex product in real arithmetic (a std::complex product is a library
// call with full C99 Annex G semantics, several times slower)
inline zc mul(zc a, zc b) {
  return zc(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}

// Split (real, imaginary) copies of a row-major complex matrix: the
// elimination loops then run on contiguous double arrays, which vectorise
// (interleaved complex rows do not).
struct Split {
  std::vector<double> re, im;
  void load(const std::vector<zc>& A) {
    re.resize(A.size());
    im.resize(A.size());
    for (size_t i = 0; i < A.size(); ++i) {
      re[i] = A[i].real();
      im[i] = A[i].imag();
    }
  }
  void store(std::vector<zc>& A) const {
    for (size_t i = 0; i < A.size(); ++i) A[i] = zc(re[i], im[i]);
  }
};

// yr + i yi -= (ar + i ai) (xr + i xi) over m entries
inline void split_axpy_neg(double* __restrict yr, double* __restrict yi, double ar, double ai,
                           const double* __restrict xr, const double* __restrict xi, int m) {
  for (int j = 0; j < m; ++j) {
    yr[j] -= ar * xr[j] - ai * xi[j];
    yi[j] -= ar * xi[j] + ai * xr[j];
  }
}

}  // namespace

// Each routine carries an AVX2/FMA clone, selected by the loader on hosts
// that have it, and a generic x86-64 one.
#define BDDA_HOST_CLONES __attribute__((target_clones("arch=haswell", "default")))

// In-place LU with partial pivoting (row interchanges recorded in piv);
// returns false if exactly singular.
BDDA_HOST_CLONES bool lu_factor(std::vector<zc>& A, std::vector<int>& piv, int n) {
  piv.resize(n);
  Split S;
  S.load(A);
  double* ar = S.re.data();
  double* ai = S.im.data();
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::hypot(ar[k * n + k], ai[k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = std::hypot(ar[i * n + k], ai[i * n + k]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    piv[k] = p;
    if (best == 0.0) return false;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        std::swap(ar[k * n + j], ar[p * n + j]);
        std::swap(ai[k * n + j], ai[p * n + j]);
      }
    const zc inv = 1.0 / zc(ar[k * n + k], ai[k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const zc l = mul(zc(ar[i * n + k], ai[i * n + k]), inv);
      ar[i * n + k] = l.real();
      ai[i * n + k] = l.imag();
      if (l == zc(0.0)) continue;
      split_axpy_neg(ar + i * n + k + 1, ai + i * n + k + 1, l.real(), l.imag(), ar + k * n + k + 1,
                     ai + k * n + k + 1, n - k - 1);
    }
  }
  S.store(A);
  return true;
}

// Solve (LU) X = B for n x m B, in place.
BDDA_HOST_CLONES void lu_solve(const std::vector<zc>& LU, const std::vector<int>& piv, int n, std::vector<zc>& B, int m) {
  Split F, X;
  F.load(LU);
  X.load(B);
  double* br = X.re.data();
  double* bi = X.im.data();
  const double* fr = F.re.data();
  const double* fi = F.im.data();
  for (int k = 0; k < n; ++k)
    if (piv[k] != k)
      for (int j = 0; j < m; ++j) {
        std::swap(br[k * m + j], br[piv[k] * m + j]);
        std::swap(bi[k * m + j], bi[piv[k] * m + j]);
      }
  for (int k = 0; k < n; ++k)
    for (int i = k + 1; i < n; ++i) {
      const double lr = fr[i * n + k], li = fi[i * n + k];
      if (lr == 0.0 && li == 0.0) continue;
      split_axpy_neg(br + i * m, bi + i * m, lr, li, br + k * m, bi + k * m, m);
    }
  for (int k = n - 1; k >= 0; --k) {
    const zc inv = 1.0 / zc(fr[k * n + k], fi[k * n + k]);
    for (int j = 0; j < m; ++j) {
      const zc v = mul(zc(br[k * m + j], bi[k * m + j]), inv);
      br[k * m + j] = v.real();
      bi[k * m + j] = v.imag();
    }
    for (int i = 0; i < k; ++i) {
      const double ur = fr[i * n + k], ui = fi[i * n + k];
      if (ur == 0.0 && ui == 0.0) continue;
      split_axpy_neg(br + i * m, bi + i * m, ur, ui, br + k * m, bi + k * m, m);
    }
  }
  X.store(B);
}

// Hermitian positive definite G = R^H R with R upper triangular; false if not PD.
BDDA_HOST_CLONES bool cholesky_upper(const std::vector<zc>& G, std::vector<zc>& R, int n) {
  R.assign((size_t)n * n, zc(0.0));
  for (int j = 0; j < n; ++j) {
    double d = G[j * n + j].real();
    for (int k = 0; k < j; ++k) d -= std::norm(R[k * n + j]);
    if (!(d > 0.0)) return false;
    const double rjj = std::sqrt(d);
    R[j * n + j] = rjj;
    for (int i = j + 1; i < n; ++i) {
      zc s = G[j * n + i];
      for (int k = 0; k < j; ++k) s -= mul(std::conj(R[k * n + j]), R[k * n + i]);
      R[j * n + i] = s / rjj;
    }
  }
  return true;
}

// Inverse of an upper triangular matrix.
BDDA_HOST_CLONES void upper_inverse(const std::vector<zc>& R, std::vector<zc>& Ri, int n) {
  Ri.assign((size_t)n * n, zc(0.0));
  for (int j = 0; j < n; ++j) {
    Ri[j * n + j] = 1.0 / R[j * n + j];
    for (int i = j - 1; i >= 0; --i) {
      zc s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += mul(R[i * n + k], Ri[k * n + j]);
      Ri[i * n + j] = -s / R[i * n + i];
    }
  }
}

// C = op(A) B  with op = transpose (no conjugate) when ta.
BDDA_HOST_CLONES void matmul(const std::vector<zc>& A, bool ta, const std::vector<zc>& B, std::vector<zc>& C, int n) {
  Split Bs, Cs;
  Bs.load(B);
  Cs.re.assign((size_t)n * n, 0.0);
  Cs.im.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const zc a = ta ? A[k * n + i] : A[i * n + k];
      if (a == zc(0.0)) continue;
      // C[i, :] += a B[k, :]
      split_axpy_neg(Cs.re.data() + i * n, Cs.im.data() + i * n, -a.real(), -a.imag(), Bs.re.data() + k * n,
                     Bs.im.data() + k * n, n);
    }
  C.resize((size_t)n * n);
  Cs.store(C);
}

double frob(const std::vector<zc>& A) {
  double s = 0.0;
  for (const zc& v : A) s += std::norm(v);
  return std::sqrt(s);
}

}  // namespace la
}  // namespace bdda
