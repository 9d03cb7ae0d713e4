// Small dense complex float64 linear algebra on the host for the L x L
// recurrence coefficients of the block-Krylov solvers: LU with partial
// pivoting (the numpy.linalg.solve / LAPACK gesv algorithm), Cholesky of a
// Hermitian positive definite Gram matrix, triangular inverse and small
// products (row-major storage), and a Householder QR of a tall block, the
// block QR's last resort.  Implemented in host_la.cpp.
#pragma once
#include <complex>
#include <vector>

namespace bdda {
namespace la {

using zc = std::complex<double>;

// In-place LU with partial pivoting (row interchanges in piv); false if
// exactly singular.
bool lu_factor(std::vector<zc>& A, std::vector<int>& piv, int n);
// Solve (LU) X = B for an n x m B, in place.
void lu_solve(const std::vector<zc>& LU, const std::vector<int>& piv, int n, std::vector<zc>& B, int m);
// Hermitian positive definite G = R^H R with R upper triangular; false if not PD.
bool cholesky_upper(const std::vector<zc>& G, std::vector<zc>& R, int n);
// Inverse of an upper triangular matrix.
void upper_inverse(const std::vector<zc>& R, std::vector<zc>& Ri, int n);
// C = op(A) B  with op = transpose (no conjugate) when ta.
void matmul(const std::vector<zc>& A, bool ta, const std::vector<zc>& B, std::vector<zc>& C, int n);
double frob(const std::vector<zc>& A);
// Reduced Householder QR of a rows x n block held as split column-major
// arrays (column l at l * rows): on return they hold Q (orthonormal
// columns) and R (n x n, row-major, upper triangular, real diagonal).
// LAPACK's zgeqrf / zungqr reflectors, as numpy.linalg.qr; it needs no
// positive definite Gram matrix, so it also factors a numerically
// rank-deficient block.
void householder_qr(std::vector<double>& re, std::vector<double>& im, long long rows, int n, std::vector<zc>& R);

}  // namespace la
}  // namespace bdda
