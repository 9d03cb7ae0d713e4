// Test driver for the host L x L algebra (csrc/host_la.cpp): reads n, m and
// complex matrices A (n x n), B (n x m) and a Hermitian positive definite G
// (n x n) as whitespace-separated real / imaginary pairs on stdin, and
// prints whether A is nonsingular, its pivots and (if so) the solution of
// A X = B, then whether G is positive definite and (if so) R of G = R^H R,
// R^-1 and R^-1 R^-1 (matmul), one number per line.
#include <cstdio>
#include <vector>

#include "host_la.h"

using bdda::la::zc;

static bool read_mat(std::vector<zc>& M, int count) {
  M.resize(count);
  for (int i = 0; i < count; ++i) {
    double re, im;
    if (std::scanf("%lf %lf", &re, &im) != 2) return false;
    M[i] = zc(re, im);
  }
  return true;
}

static void print_mat(const std::vector<zc>& M) {
  for (const zc& v : M) std::printf("%.17g %.17g\n", v.real(), v.imag());
}

int main() {
  int n, m;
  if (std::scanf("%d %d", &n, &m) != 2) return 2;
  std::vector<zc> A, B, G;
  if (!read_mat(A, n * n) || !read_mat(B, n * m) || !read_mat(G, n * n)) return 2;
  std::vector<int> piv;
  const bool ok = bdda::la::lu_factor(A, piv, n);
  std::printf("%d\n", ok ? 1 : 0);
  for (int p : piv) std::printf("%d\n", p);
  if (ok) {
    bdda::la::lu_solve(A, piv, n, B, m);
    print_mat(B);
  }
  std::vector<zc> R, Ri, C;
  const bool pd = bdda::la::cholesky_upper(G, R, n);
  std::printf("%d\n", pd ? 1 : 0);
  if (pd) {
    bdda::la::upper_inverse(R, Ri, n);
    bdda::la::matmul(Ri, false, Ri, C, n);
    print_mat(R);
    print_mat(Ri);
    print_mat(C);
  }
  return 0;
}
