// Small dense complex float64 linear algebra on the host for the L x L
// recurrence coefficients of the block-Krylov solvers (L <= 128): LU with
// partial pivoting (the numpy.linalg.solve / LAPACK gesv algorithm),
// Cholesky of a Hermitian positive definite Gram matrix, triangular inverse
// and small products.  Row-major storage.
#pragma once
#include <cmath>
#include <complex>
#include <vector>

namespace bdda {
namespace la {

using zc = std::complex<double>;

// In-place LU with partial pivoting; returns false if exactly singular.
inline bool lu_factor(std::vector<zc>& A, std::vector<int>& piv, int n) {
  piv.resize(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::abs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = std::abs(A[i * n + k]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    piv[k] = p;
    if (best == 0.0) return false;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
    const zc inv = 1.0 / A[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const zc l = A[i * n + k] * inv;
      A[i * n + k] = l;
      if (l == zc(0.0))
        continue;
      zc* ri = &A[i * n];
      const zc* rk = &A[k * n];
      for (int j = k + 1; j < n; ++j) ri[j] -= l * rk[j];
    }
  }
  return true;
}

// Solve (LU) X = B for n x m B, in place.
inline void lu_solve(const std::vector<zc>& LU, const std::vector<int>& piv, int n, std::vector<zc>& B, int m) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k)
      for (int j = 0; j < m; ++j) std::swap(B[k * m + j], B[piv[k] * m + j]);
  for (int k = 0; k < n; ++k)
    for (int i = k + 1; i < n; ++i) {
      const zc l = LU[i * n + k];
      if (l == zc(0.0)) continue;
      for (int j = 0; j < m; ++j) B[i * m + j] -= l * B[k * m + j];
    }
  for (int k = n - 1; k >= 0; --k) {
    const zc inv = 1.0 / LU[k * n + k];
    for (int j = 0; j < m; ++j) B[k * m + j] *= inv;
    for (int i = 0; i < k; ++i) {
      const zc u = LU[i * n + k];
      if (u == zc(0.0)) continue;
      for (int j = 0; j < m; ++j) B[i * m + j] -= u * B[k * m + j];
    }
  }
}

// Hermitian positive definite G = R^H R with R upper triangular; false if not PD.
inline bool cholesky_upper(const std::vector<zc>& G, std::vector<zc>& R, int n) {
  R.assign((size_t)n * n, zc(0.0));
  for (int j = 0; j < n; ++j) {
    double d = G[j * n + j].real();
    for (int k = 0; k < j; ++k) d -= std::norm(R[k * n + j]);
    if (!(d > 0.0)) return false;
    const double rjj = std::sqrt(d);
    R[j * n + j] = rjj;
    for (int i = j + 1; i < n; ++i) {
      zc s = G[j * n + i];
      for (int k = 0; k < j; ++k) s -= std::conj(R[k * n + j]) * R[k * n + i];
      R[j * n + i] = s / rjj;
    }
  }
  return true;
}

// Inverse of an upper triangular matrix.
inline void upper_inverse(const std::vector<zc>& R, std::vector<zc>& Ri, int n) {
  Ri.assign((size_t)n * n, zc(0.0));
  for (int j = 0; j < n; ++j) {
    Ri[j * n + j] = 1.0 / R[j * n + j];
    for (int i = j - 1; i >= 0; --i) {
      zc s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += R[i * n + k] * Ri[k * n + j];
      Ri[i * n + j] = -s / R[i * n + i];
    }
  }
}

// C = op(A) B  with op = transpose (no conjugate) when ta.
inline void matmul(const std::vector<zc>& A, bool ta, const std::vector<zc>& B, std::vector<zc>& C, int n) {
  C.assign((size_t)n * n, zc(0.0));
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const zc a = ta ? A[k * n + i] : A[i * n + k];
      if (a == zc(0.0)) continue;
      for (int j = 0; j < n; ++j) C[i * n + j] += a * B[k * n + j];
    }
}

inline double frob(const std::vector<zc>& A) {
  double s = 0.0;
  for (const zc& v : A) s += std::norm(v);
  return std::sqrt(s);
}

}  // namespace la
}  // namespace bdda
