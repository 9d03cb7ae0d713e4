// Test driver for the host Householder QR (csrc/host_la.cpp): reads rows, n
// and a rows x n complex matrix (row-major, real / imaginary pairs) on
// stdin; prints Q (rows x n, row-major) and R (n x n), one number per line.
#include <cstdio>
#include <vector>

#include "host_la.h"

using bdda::la::zc;

int main() {
  long long rows;
  int n;
  if (std::scanf("%lld %d", &rows, &n) != 2) return 2;
  std::vector<double> re((size_t)rows * n), im((size_t)rows * n);
  for (long long r = 0; r < rows; ++r)
    for (int l = 0; l < n; ++l)
      if (std::scanf("%lf %lf", &re[(size_t)l * rows + r], &im[(size_t)l * rows + r]) != 2) return 2;
  std::vector<zc> R;
  bdda::la::householder_qr(re, im, rows, n, R);
  for (long long r = 0; r < rows; ++r)
    for (int l = 0; l < n; ++l) std::printf("%.17g %.17g\n", re[(size_t)l * rows + r], im[(size_t)l * rows + r]);
  for (const zc& v : R) std::printf("%.17g %.17g\n", v.real(), v.imag());
  return 0;
}
