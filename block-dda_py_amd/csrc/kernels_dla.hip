// Dense L x L complex float64 algebra on the device (L <= 64) for the
// recurrence coefficients of the block-Krylov solvers: LU with partial
// pivoting and the solve that follows it, Cholesky of a Hermitian Gram
// matrix with its upper-triangular inverse, small products and the scalar
// reductions of block-BiCGSTAB.  Each routine is one workgroup working in
// LDS, so a solver iteration keeps its coefficients on the device and waits
// on the host once (for the convergence test) instead of after every Gram
// matrix.  The algorithms and their orders follow the host routines
// (host_la.cpp): the pivot is the first entry of largest |Re| + |Im| of the
// column (LAPACK zgetrf through izamax, as numpy.linalg.solve in the
// reference), Cholesky is right-looking a row at a time, the triangular
// inverse runs a row at a time from the bottom.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"

namespace bdda {

namespace {

constexpr int kT = 256;                 // threads per workgroup
constexpr int kP = kDlaMax + 1;         // LDS row pitch (odd: rows start on distinct banks)

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zrecip(double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2(b.x / d, -b.y / d);
}
// y -= a x
__device__ __forceinline__ void zsubmul(double2& y, double2 a, double2 x) {
  y.x -= a.x * x.x - a.y * x.y;
  y.y -= a.x * x.y + a.y * x.x;
}

// Fixed-order sum over the workgroup (the same tree every call).
__device__ double wg_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < kT / 64; ++i) t += sh[i];
  __syncthreads();
  return t;
}

// Forward elimination with partial pivoting of a (n x n, in LDS) carried
// along to b (n x m, in LDS); back substitution into b.  On an exactly
// singular pivot *sing is set and the elimination stops (b is left as is).
__device__ void lu_eliminate_solve(double2* a, double2* b, int n, int m, int* piv, int* sing, double* shv,
                                   int* shi) {
  const int tid = threadIdx.x;
  for (int k = 0; k < n; ++k) {
    // pivot: first row of largest |Re| + |Im| in column k (wave 0)
    if (tid < 64) {
      double best = -1.0;
      int bi = n;
      for (int i = k + tid; i < n; i += 64) {
        const double2 v = a[i * kP + k];
        const double s = fabs(v.x) + fabs(v.y);
        if (s > best) {
          best = s;
          bi = i;
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_down(best, o, 64);
        const int oi = __shfl_down(bi, o, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (tid == 0) {
        shv[0] = best;
        shi[0] = bi;
      }
    }
    __syncthreads();
    const int p = shi[0];
    if (!(shv[0] > 0.0)) {
      if (tid == 0) *sing = 1;
      return;
    }
    if (tid == 0) piv[k] = p;
    if (p != k) {
      for (int j = tid; j < n; j += kT) {
        const double2 t = a[k * kP + j];
        a[k * kP + j] = a[p * kP + j];
        a[p * kP + j] = t;
      }
      for (int j = tid; j < m; j += kT) {
        const double2 t = b[k * kP + j];
        b[k * kP + j] = b[p * kP + j];
        b[p * kP + j] = t;
      }
    }
    __syncthreads();
    const double2 inv = zrecip(a[k * kP + k]);
    for (int i = k + 1 + tid; i < n; i += kT) a[i * kP + k] = zmul(a[i * kP + k], inv);
    __syncthreads();
    // trailing update on an 8 x 32 thread grid (rows by columns of [A | B])
    const int ty = tid >> 5, tx = tid & 31;
    for (int i = k + 1 + ty; i < n; i += 8) {
      const double2 l = a[i * kP + k];
      for (int j = k + 1 + tx; j < n; j += 32) zsubmul(a[i * kP + j], l, a[k * kP + j]);
      for (int j = tx; j < m; j += 32) zsubmul(b[i * kP + j], l, b[k * kP + j]);
    }
    __syncthreads();
  }
  // back substitution: b <- U^-1 b, one row of U at a time from the bottom
  const int ty = tid >> 5, tx = tid & 31;
  for (int k = n - 1; k >= 0; --k) {
    const double2 inv = zrecip(a[k * kP + k]);
    for (int j = tid; j < m; j += kT) b[k * kP + j] = zmul(b[k * kP + j], inv);
    __syncthreads();
    for (int i = ty; i < k; i += 8) {
      const double2 u = a[i * kP + k];
      for (int j = tx; j < m; j += 32) zsubmul(b[i * kP + j], u, b[k * kP + j]);
    }
    __syncthreads();
  }
}

// X = A^-1 op(B) for an n x n A and n x m B (op: negation when neg_b), with
// Xneg = -X (nullable), and the LU factors and pivots of A (nullable) for a
// later dla_lu_resolve.  status bit DLA_SINGULAR on an exactly singular A,
// X and Xneg then zero.
__global__ void __launch_bounds__(kT) k_dla_lu_solve(const double2* __restrict__ A, const double2* __restrict__ B,
                                                    int n, int m, int neg_b, double2* X, double2* Xneg, double2* LU,
                                                    int* piv_out, int* status) {
  extern __shared__ __align__(16) unsigned char dla_smem[];
  double2* a = reinterpret_cast<double2*>(dla_smem);
  double2* b = a + kDlaMax * kP;
  __shared__ double shv[1];
  __shared__ int shi[1], sing[1];
  __shared__ int piv[kDlaMax];
  const int tid = threadIdx.x;
  if (tid == 0) sing[0] = 0;
  for (int idx = tid; idx < n * n; idx += kT) a[(idx / n) * kP + idx % n] = A[idx];
  for (int idx = tid; idx < n * m; idx += kT) {
    double2 v = B[idx];
    if (neg_b) v = make_double2(-v.x, -v.y);
    b[(idx / m) * kP + idx % m] = v;
  }
  __syncthreads();
  // once an earlier coefficient system of the iteration was singular, every
  // later one yields zero coefficients too, so that no update changes
  // anything until the host sees the status
  const bool prior = (*status & DLA_SINGULAR) != 0;
  if (!prior) lu_eliminate_solve(a, b, n, m, piv, sing, shv, shi);
  __syncthreads();
  const bool bad = prior || sing[0] != 0;
  if (bad && tid == 0) atomicOr(status, DLA_SINGULAR);
  for (int idx = tid; idx < n * m; idx += kT) {
    const double2 v = bad ? make_double2(0.0, 0.0) : b[(idx / m) * kP + idx % m];
    X[idx] = v;
    if (Xneg) Xneg[idx] = make_double2(-v.x, -v.y);
  }
  if (LU) {
    for (int idx = tid; idx < n * n; idx += kT) LU[idx] = a[(idx / n) * kP + idx % n];
    for (int i = tid; i < n; i += kT) piv_out[i] = piv[i];
  }
}

// X = A^-1 op(B) from the factors of k_dla_lu_solve (the host's lu_solve:
// row interchanges, unit-lower forward substitution, back substitution).
__global__ void __launch_bounds__(kT) k_dla_lu_resolve(const double2* __restrict__ LU, const int* __restrict__ piv,
                                                      const double2* __restrict__ B, int n, int m, int neg_b,
                                                      double2* X, const int* status) {
  extern __shared__ __align__(16) unsigned char dla_smem[];
  double2* a = reinterpret_cast<double2*>(dla_smem);
  double2* b = a + kDlaMax * kP;
  const int tid = threadIdx.x;
  const bool bad = (*status & DLA_SINGULAR) != 0;
  for (int idx = tid; idx < n * n; idx += kT) a[(idx / n) * kP + idx % n] = LU[idx];
  for (int idx = tid; idx < n * m; idx += kT) {
    double2 v = B[idx];
    if (neg_b) v = make_double2(-v.x, -v.y);
    b[(idx / m) * kP + idx % m] = v;
  }
  __syncthreads();
  if (!bad) {
    for (int k = 0; k < n; ++k) {
      const int p = piv[k];
      if (p != k)
        for (int j = tid; j < m; j += kT) {
          const double2 t = b[k * kP + j];
          b[k * kP + j] = b[p * kP + j];
          b[p * kP + j] = t;
        }
      __syncthreads();
    }
    const int ty = tid >> 5, tx = tid & 31;
    for (int k = 0; k < n; ++k) {
      for (int i = k + 1 + ty; i < n; i += 8) {
        const double2 l = a[i * kP + k];
        for (int j = tx; j < m; j += 32) zsubmul(b[i * kP + j], l, b[k * kP + j]);
      }
      __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
      const double2 inv = zrecip(a[k * kP + k]);
      for (int j = tid; j < m; j += kT) b[k * kP + j] = zmul(b[k * kP + j], inv);
      __syncthreads();
      for (int i = ty; i < k; i += 8) {
        const double2 u = a[i * kP + k];
        for (int j = tx; j < m; j += 32) zsubmul(b[i * kP + j], u, b[k * kP + j]);
      }
      __syncthreads();
    }
  }
  for (int idx = tid; idx < n * m; idx += kT) X[idx] = bad ? make_double2(0.0, 0.0) : b[(idx / m) * kP + idx % m];
}

// Right-looking Cholesky G = R^H R of the n x n matrix in a (LDS, upper
// part used); false if a pivot is not positive.
__device__ bool chol_upper(double2* a, int n) {
  const int tid = threadIdx.x;
  for (int j = 0; j < n; ++j) {
    const double d = a[j * kP + j].x;
    if (!(d > 0.0)) return false;
    const double rjj = sqrt(d), inv = 1.0 / rjj;
    __syncthreads();
    if (tid == 0) a[j * kP + j] = make_double2(rjj, 0.0);
    for (int l = j + 1 + tid; l < n; l += kT) {
      a[j * kP + l].x *= inv;
      a[j * kP + l].y *= inv;
    }
    __syncthreads();
    // rows i > j from column i on: A[i][l] -= conj(R[j][i]) R[j][l]
    const int ty = tid >> 5, tx = tid & 31;
    for (int i = j + 1 + ty; i < n; i += 8) {
      const double2 c = make_double2(a[j * kP + i].x, -a[j * kP + i].y);
      for (int l = i + tx; l < n; l += 32) zsubmul(a[i * kP + l], c, a[j * kP + l]);
    }
    __syncthreads();
  }
  return true;
}

// R (upper, zeros below) and Ri = R^-1 with R^H R = G (+ s I), the shift
// s = 11 (rows n + n (n + 1)) 2^-53 tr(G) applied only when `shift` is set
// and the plain factorisation fails (status DLA_SHIFTED); status DLA_NOT_PD
// if no factorisation succeeds (R, Ri zero).
__global__ void __launch_bounds__(kT) k_dla_chol_inv(const double2* __restrict__ G, int n, double rows, int shift,
                                                    double2* R, double2* Ri, int* status) {
  extern __shared__ __align__(16) unsigned char dla_smem[];
  double2* a = reinterpret_cast<double2*>(dla_smem);
  double2* x = a + kDlaMax * kP;
  __shared__ double sh[kT / 64];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < n * n; idx += kT) a[(idx / n) * kP + idx % n] = G[idx];
  __syncthreads();
  bool ok = chol_upper(a, n);
  __syncthreads();
  if (!ok && shift) {
    double tr = 0.0;
    for (int i = tid; i < n; i += kT) tr += G[i * n + i].x;
    tr = wg_sum(tr, sh);
    const double s = 11.0 * (rows * n + (double)n * (n + 1)) * 0x1p-53 * tr;
    for (int idx = tid; idx < n * n; idx += kT) {
      const int i = idx / n, j = idx % n;
      double2 v = G[idx];
      if (i == j) v.x += s;
      a[i * kP + j] = v;
    }
    __syncthreads();
    ok = chol_upper(a, n);
    __syncthreads();
    if (tid == 0) atomicOr(status, ok ? DLA_SHIFTED : DLA_NOT_PD);
  } else if (!ok && tid == 0) {
    atomicOr(status, DLA_NOT_PD);
  }
  // inverse of the upper factor, a row at a time from the bottom:
  // X[i][l] = -(1 / R_ii) sum_{i < k <= l} R_ik X[k][l], X[i][i] = 1 / R_ii
  for (int idx = tid; idx < n * n; idx += kT) x[(idx / n) * kP + idx % n] = make_double2(0.0, 0.0);
  __syncthreads();
  if (ok) {
    for (int i = n - 1; i >= 0; --i) {
      const double2 inv = zrecip(a[i * kP + i]);
      for (int l = i + 1 + tid; l < n; l += kT) {
        double2 s = make_double2(0.0, 0.0);
        for (int k = i + 1; k <= l; ++k) zsubmul(s, make_double2(-a[i * kP + k].x, -a[i * kP + k].y), x[k * kP + l]);
        x[i * kP + l] = zmul(s, make_double2(-inv.x, -inv.y));
      }
      if (tid == 0) x[i * kP + i] = inv;
      __syncthreads();
    }
  }
  for (int idx = tid; idx < n * n; idx += kT) {
    const int i = idx / n, j = idx % n;
    R[idx] = (ok && j >= i) ? a[i * kP + j] : make_double2(0.0, 0.0);
    Ri[idx] = ok ? x[i * kP + j] : make_double2(0.0, 0.0);
  }
}

// C = op(A) B for n x n matrices (op = transpose when ta); C must not alias.
__global__ void __launch_bounds__(kT) k_dla_matmul(const double2* __restrict__ A, int ta,
                                                  const double2* __restrict__ B, double2* __restrict__ C, int n) {
  for (int idx = threadIdx.x; idx < n * n; idx += kT) {
    const int i = idx / n, j = idx % n;
    double2 s = make_double2(0.0, 0.0);
    for (int k = 0; k < n; ++k) {
      const double2 av = ta ? A[k * n + i] : A[i * n + k];
      s.x += av.x * B[k * n + j].x - av.y * B[k * n + j].y;
      s.y += av.x * B[k * n + j].y + av.y * B[k * n + j].x;
    }
    C[idx] = s;
  }
}

// out[0] = ||A||_F / scale (n x n), in a fixed-order workgroup sum.
__global__ void __launch_bounds__(kT) k_dla_frob(const double2* __restrict__ A, int n, double scale,
                                                double* out) {
  __shared__ double sh[kT / 64];
  double s = 0.0;
  for (int idx = threadIdx.x; idx < n * n; idx += kT) s += A[idx].x * A[idx].x + A[idx].y * A[idx].y;
  s = wg_sum(s, sh);
  if (threadIdx.x == 0) out[0] = sqrt(s) / scale;
}

// Block-BiCGSTAB's xi = tr(Z^H T) / tr(Z^H Z) from the dot partials of
// k_dot (sum conj(T) Z and |Z|^2, summed in workgroup order as the host
// does): sc[0] = -xi, sc[1] = xi (zero when the iteration's L x L system was
// singular, so that its updates change nothing).
constexpr int kMaxParts = 1024;

__global__ void __launch_bounds__(kT) k_dla_xi(const double2* __restrict__ pdot, const double* __restrict__ pnrm,
                                              int np, const int* status, double2* sc) {
  __shared__ double2 sd[kMaxParts];
  __shared__ double sn[kMaxParts];
  for (int i = threadIdx.x; i < np; i += kT) {
    sd[i] = pdot[i];
    sn[i] = pnrm[i];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double tx = 0.0, ty = 0.0, z = 0.0;
  for (int i = 0; i < np; ++i) {
    tx += sd[i].x;
    ty += sd[i].y;
    z += sn[i];
  }
  double2 xi = make_double2(tx / z, -ty / z);  // conj(tz) / |Z|^2
  if (*status & DLA_SINGULAR) xi = make_double2(0.0, 0.0);
  sc[0] = make_double2(-xi.x, -xi.y);
  sc[1] = xi;
}

// out[0] = sqrt(sum of the norm partials) / scale, summed in workgroup
// order as the host does.
__global__ void __launch_bounds__(kT) k_dla_norm_err(const double* __restrict__ pnrm, int np, double scale,
                                                      double* out) {
  __shared__ double sn[kMaxParts];
  for (int i = threadIdx.x; i < np; i += kT) sn[i] = pnrm[i];
  __syncthreads();
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < np; ++i) s += sn[i];
  out[0] = sqrt(s) / scale;
}

constexpr size_t kDlaLds = 2 * (size_t)kDlaMax * kP * sizeof(double2);

void set_lds_attr(const void* f) {
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDlaLds);
}

std::once_flag g_dla_attr;
void dla_attrs() {
  std::call_once(g_dla_attr, [] {
    set_lds_attr(reinterpret_cast<const void*>(&k_dla_lu_solve));
    set_lds_attr(reinterpret_cast<const void*>(&k_dla_lu_resolve));
    set_lds_attr(reinterpret_cast<const void*>(&k_dla_chol_inv));
  });
}

}  // namespace

void dla_lu_solve(const double2* A, const double2* B, int n, int m, bool neg_b, double2* X, double2* Xneg,
                  double2* LU, int* piv, int* status, hipStream_t s) {
  dla_attrs();
  hipLaunchKernelGGL(k_dla_lu_solve, dim3(1), dim3(kT), kDlaLds, s, A, B, n, m, (int)neg_b, X, Xneg, LU, piv,
                     status);
}
void dla_lu_resolve(const double2* LU, const int* piv, const double2* B, int n, int m, bool neg_b, double2* X,
                    const int* status, hipStream_t s) {
  dla_attrs();
  hipLaunchKernelGGL(k_dla_lu_resolve, dim3(1), dim3(kT), kDlaLds, s, LU, piv, B, n, m, (int)neg_b, X, status);
}
void dla_chol_inv(const double2* G, int n, double rows, bool shift, double2* R, double2* Ri, int* status,
                  hipStream_t s) {
  dla_attrs();
  hipLaunchKernelGGL(k_dla_chol_inv, dim3(1), dim3(kT), kDlaLds, s, G, n, rows, (int)shift, R, Ri, status);
}
void dla_matmul(const double2* A, bool ta, const double2* B, double2* C, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_dla_matmul, dim3(1), dim3(kT), 0, s, A, (int)ta, B, C, n);
}
void dla_frob(const double2* A, int n, double scale, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_dla_frob, dim3(1), dim3(kT), 0, s, A, n, scale, out);
}
void dla_xi(const double2* pdot, const double* pnrm, int np, const int* status, double2* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_dla_xi, dim3(1), dim3(kT), 0, s, pdot, pnrm, np, status, sc);
}
void dla_norm_err(const double* pnrm, int np, double scale, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_dla_norm_err, dim3(1), dim3(kT), 0, s, pnrm, np, scale, out);
}

}  // namespace bdda
