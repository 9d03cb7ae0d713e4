// Small dense complex float64 linear algebra on the host for the L x L
// recurrence coefficients of the block-Krylov solvers (declarations and
// conventions in host_la.h).  Plain C++ (built with the host compiler), so
// each routine can carry AVX-512 and AVX2/FMA clones that the loader selects
// at run time, with a generic x86-64 fallback.
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

// Each routine carries AVX-512 and AVX2/FMA clones, selected by the loader
// on hosts that have them, and a generic x86-64 one.
#define BDDA_HOST_CLONES __attribute__((target_clones("avx512f", "arch=haswell", "default")))

// In-place LU with partial pivoting (row interchanges recorded in piv);
// returns false if exactly singular.  The pivot is the first entry of
// largest |Re| + |Im| in the column, LAPACK's choice (zgetrf through
// izamax), which np.linalg.solve in the reference makes.
BDDA_HOST_CLONES bool lu_factor(std::vector<zc>& A, std::vector<int>& piv, int n) {
  piv.resize(n);
  Split S;
  S.load(A);
  double* ar = S.re.data();
  double* ai = S.im.data();
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(ar[k * n + k]) + std::fabs(ai[k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = std::fabs(ar[i * n + k]) + std::fabs(ai[i * n + k]);
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
// Right-looking, one row at a time: row j of R is row j of the updated
// matrix over r_jj, then every later row i takes its update
// -conj(R[j][i]) R[j][i..n) (contiguous, vectorises).  Each entry still
// accumulates its products in the order j = 0, 1, ...
BDDA_HOST_CLONES bool cholesky_upper(const std::vector<zc>& G, std::vector<zc>& R, int n) {
  Split A;
  A.load(G);
  double* ar = A.re.data();
  double* ai = A.im.data();
  for (int j = 0; j < n; ++j) {
    const double d = ar[j * n + j];
    if (!(d > 0.0)) return false;
    const double rjj = std::sqrt(d), inv = 1.0 / rjj;
    ar[j * n + j] = rjj;
    ai[j * n + j] = 0.0;
    for (int l = j + 1; l < n; ++l) {
      ar[j * n + l] *= inv;
      ai[j * n + l] *= inv;
    }
    for (int i = j + 1; i < n; ++i) {
      // row i from column i on: A[i][l] -= conj(R[j][i]) R[j][l]
      const double cr = ar[j * n + i], ci = -ai[j * n + i];
      split_axpy_neg(ar + i * n + i, ai + i * n + i, cr, ci, ar + j * n + i, ai + j * n + i, n - i);
    }
  }
  R.assign((size_t)n * n, zc(0.0));
  for (int j = 0; j < n; ++j)
    for (int l = j; l < n; ++l) R[j * n + l] = zc(ar[j * n + l], ai[j * n + l]);
  return true;
}

// Inverse of an upper triangular matrix, a row at a time from the bottom:
// row i of R^-1 is -(1 / R_ii) sum_{k > i} R_ik row k, plus 1 / R_ii on the
// diagonal (the row updates are contiguous and vectorise).
BDDA_HOST_CLONES void upper_inverse(const std::vector<zc>& R, std::vector<zc>& Ri, int n) {
  Split F, X;
  F.load(R);
  X.re.assign((size_t)n * n, 0.0);
  X.im.assign((size_t)n * n, 0.0);
  const double* fr = F.re.data();
  const double* fi = F.im.data();
  double* xr = X.re.data();
  double* xi = X.im.data();
  for (int i = n - 1; i >= 0; --i) {
    for (int k = i + 1; k < n; ++k) {
      // X[i][k..n) += R[i][k] X[k][k..n)
      split_axpy_neg(xr + i * n + k, xi + i * n + k, -fr[i * n + k], -fi[i * n + k], xr + k * n + k,
                     xi + k * n + k, n - k);
    }
    const zc inv = 1.0 / zc(fr[i * n + i], fi[i * n + i]);
    for (int l = i + 1; l < n; ++l) {
      const zc v = mul(zc(xr[i * n + l], xi[i * n + l]), -inv);
      xr[i * n + l] = v.real();
      xi[i * n + l] = v.imag();
    }
    xr[i * n + i] = inv.real();
    xi[i * n + i] = inv.imag();
  }
  Ri.resize((size_t)n * n);
  X.store(Ri);
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

namespace {
// w = v^H y over entries [j, rows) with v[j] = 1 (the reflector's head)
inline zc reflector_dot(const double* vr, const double* vi, const double* yr, const double* yi, long long j,
                        long long rows) {
  double wr = yr[j], wi = yi[j];
  for (long long r = j + 1; r < rows; ++r) {
    wr += vr[r] * yr[r] + vi[r] * yi[r];
    wi += vr[r] * yi[r] - vi[r] * yr[r];
  }
  return zc(wr, wi);
}
// y -= f v over entries [j, rows) with v[j] = 1
inline void reflector_axpy(const double* vr, const double* vi, double* yr, double* yi, zc f, long long j,
                           long long rows) {
  yr[j] -= f.real();
  yi[j] -= f.imag();
  split_axpy_neg(yr + j + 1, yi + j + 1, f.real(), f.imag(), vr + j + 1, vi + j + 1, (int)(rows - j - 1));
}
}  // namespace

// zgeqr2: for each column j the reflector H_j = I - tau_j v v^H (zlarfg:
// beta = -sign(Re a_jj) |a_j[j:]|, tau_j = (beta - a_jj) / beta, v = a_j[j:] /
// (a_jj - beta) with v_j = 1) maps a_j[j:] to beta e_j; H_j^H is applied to
// the later columns.  zung2r then forms Q = H_0 ... H_{n-1} [I; 0] from the
// last reflector back.
BDDA_HOST_CLONES void householder_qr(std::vector<double>& re, std::vector<double>& im, long long rows, int n,
                                     std::vector<zc>& R) {
  std::vector<zc> tau(n, zc(0.0));
  for (int j = 0; j < n; ++j) {
    double* vr = re.data() + (size_t)j * rows;
    double* vi = im.data() + (size_t)j * rows;
    const zc alpha(vr[j], vi[j]);
    double s = 0.0;
    for (long long r = j + 1; r < rows; ++r) s += vr[r] * vr[r] + vi[r] * vi[r];
    if (s == 0.0 && alpha.imag() == 0.0) continue;  // H_j = I
    double beta = std::sqrt(std::norm(alpha) + s);
    if (alpha.real() >= 0.0) beta = -beta;
    tau[j] = zc((beta - alpha.real()) / beta, -alpha.imag() / beta);
    const zc scale = 1.0 / (alpha - beta);
    for (long long r = j + 1; r < rows; ++r) {
      const zc v = mul(zc(vr[r], vi[r]), scale);
      vr[r] = v.real();
      vi[r] = v.imag();
    }
    vr[j] = beta;
    vi[j] = 0.0;
    for (int k = j + 1; k < n; ++k) {
      double* yr = re.data() + (size_t)k * rows;
      double* yi = im.data() + (size_t)k * rows;
      const zc f = mul(std::conj(tau[j]), reflector_dot(vr, vi, yr, yi, j, rows));
      reflector_axpy(vr, vi, yr, yi, f, j, rows);
    }
  }
  R.assign((size_t)n * n, zc(0.0));
  for (int i = 0; i < n; ++i)
    for (int l = i; l < n; ++l) R[(size_t)i * n + l] = zc(re[(size_t)l * rows + i], im[(size_t)l * rows + i]);
  for (int j = n - 1; j >= 0; --j) {
    double* vr = re.data() + (size_t)j * rows;
    double* vi = im.data() + (size_t)j * rows;
    for (int k = j + 1; k < n; ++k) {
      double* yr = re.data() + (size_t)k * rows;
      double* yi = im.data() + (size_t)k * rows;
      const zc f = mul(tau[j], reflector_dot(vr, vi, yr, yi, j, rows));
      reflector_axpy(vr, vi, yr, yi, f, j, rows);
    }
    for (long long r = j + 1; r < rows; ++r) {
      const zc v = mul(zc(vr[r], vi[r]), -tau[j]);
      vr[r] = v.real();
      vi[r] = v.imag();
    }
    vr[j] = 1.0 - tau[j].real();
    vi[j] = -tau[j].imag();
    for (long long r = 0; r < j; ++r) vr[r] = vi[r] = 0.0;
  }
}

}  // namespace la
}  // namespace bdda
