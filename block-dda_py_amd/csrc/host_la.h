// Small dense complex float64 linear algebra on the host for the L x L
// recurrence coefficients of the block-Krylov solvers (L <= 128): LU with
// partial pivoting (the numpy.linalg.solve / LAPACK gesv algorithm),
// Cholesky of a Hermitian positive definite Gram matrix, triangular inverse
// and small products.  Row-major storage; implemented in host_la.cpp.
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

}  // namespace la
}  // namespace bdda
