// Tall-skinny block kernels of the block-Krylov solvers, the right-hand side
// and observable reductions, and the GRE voxeliser, for gfx950.
//
// Krylov blocks are (rows, L) row-major with rows = 3N (row = 3m + c) and the
// L orientation columns innermost, so a wavefront reads 64 consecutive
// columns of one row.  Small L x L coefficient matrices are complex float64.
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "internal.h"

namespace bdda {

template <typename V> struct RealOf;
template <> struct RealOf<double2> { using T = double; };
template <> struct RealOf<float2> { using T = float; };

__device__ __forceinline__ double2 d2(float2 v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ double2 d2(double2 v) { return v; }
template <typename V> __device__ __forceinline__ V from_d2(double2 v);
template <> __device__ __forceinline__ double2 from_d2<double2>(double2 v) { return v; }
template <> __device__ __forceinline__ float2 from_d2<float2>(double2 v) {
  return make_float2((float)v.x, (float)v.y);
}
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// --------------------------------------------------------------------- Gram
// partial[wg][tile](64x64) = sum_{rows of wg} op(A[i][jt*64+a]) * B[i][lt*64+b]
// Row ranges (parts) per Gram: 512 up to L = 128; beyond, fewer, so that
// the partial tiles (parts x ceil(L/64)^2 x 64 x 64) stay near 64 MB while
// the grid keeps at least 512 blocks; and at least kGramMinRows rows per
// part, so that on small blocks the partial tiles (and their reduction) do
// not outweigh the rows they sum.
constexpr int kGramParts = 512;
constexpr long long kGramMinRows = 128;
int gram_parts(int L, long long rows) {
  const int nt = (L + 63) / 64;
  const int most = nt <= 2 ? kGramParts : std::max(8, 4 * kGramParts / (nt * nt));
  return (int)std::max(1LL, std::min<long long>(most, rows / kGramMinRows));
}

// A block owns a contiguous range of rows and one TL x TL output tile
// (TL = 64, 32 or 16: the smallest that holds L, so that small L do not
// multiply padding).  Its 256 threads form NG = 256 / (TL/4)^2 groups, each
// with a 4 x 4 register tile per thread (rows ty + TL/4 i, columns
// tx + TL/4 j: lanes walk consecutive LDS words, so reads are
// conflict-free); group g takes the rows g, g + NG, ... of every 16-row slab.
// The next slab is loaded into registers while the current one is
// multiplied.  At the end the groups' tiles are summed in group order
// through LDS and the block writes one partial tile.
template <typename V, bool CONJ, int TL>
__global__ void __launch_bounds__(256) k_gram(const V* __restrict__ A, const V* __restrict__ B, long long rows,
                                              int L, double2* __restrict__ partial) {
  constexpr int TQ = TL / 4, TPG = TQ * TQ, NG = 256 / TPG;
  constexpr int PER = (16 * TL) / 256;  // slab entries per thread
  static_assert(NG == 1 || TL * TL <= 32 * TL, "the group sum reuses the slab buffers");
  __shared__ double2 sh[32 * TL];
  double2 (*As)[TL] = reinterpret_cast<double2 (*)[TL]>(sh);
  double2 (*Bs)[TL] = reinterpret_cast<double2 (*)[TL]>(sh + 16 * TL);
  const int tid = threadIdx.x, g = tid / TPG, t = tid % TPG, tx = t % TQ, ty = t / TQ;
  const int ntile = (L + TL - 1) / TL;
  const int jt = blockIdx.y / ntile, lt = blockIdx.y % ntile;
  const long long chunk = (rows + gridDim.x - 1) / gridDim.x;
  const long long r0 = blockIdx.x * chunk;
  const long long r1 = min(rows, r0 + chunk);
  double2 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = make_double2(0.0, 0.0);
  // this thread's share of a slab: entries tid + 256 u (u < PER) of [16][TL]
  double2 pa[PER], pb[PER];
  auto fetch = [&](long long rs) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int idx = tid + 256 * u, rr = idx / TL, cc = idx % TL;
      const long long row = rs + rr;
      const int ja = jt * TL + cc, lb = lt * TL + cc;
      pa[u] = (row < r1 && ja < L) ? d2(A[row * L + ja]) : make_double2(0.0, 0.0);
      pb[u] = (row < r1 && lb < L) ? d2(B[row * L + lb]) : make_double2(0.0, 0.0);
      if (CONJ) pa[u].y = -pa[u].y;
    }
  };
  if (r0 < r1) fetch(r0);
  for (long long rs = r0; rs < r1; rs += 16) {
    __syncthreads();  // the previous slab has been consumed
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int idx = tid + 256 * u;
      As[idx / TL][idx % TL] = pa[u];
      Bs[idx / TL][idx % TL] = pb[u];
    }
    __syncthreads();
    if (rs + 16 < r1) fetch(rs + 16);  // in flight during the products
#pragma unroll
    for (int rr = g; rr < 16; rr += NG) {
      double2 a[4], b[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = As[rr][ty + TQ * q];
        b[q] = Bs[rr][tx + TQ * q];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          acc[i][j].x += a[i].x * b[j].x - a[i].y * b[j].y;
          acc[i][j].y += a[i].x * b[j].y + a[i].y * b[j].x;
        }
    }
  }
  double2* out = partial + ((long long)blockIdx.x * gridDim.y + blockIdx.y) * (TL * TL);
  if constexpr (NG == 1) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) out[(ty + TQ * i) * TL + tx + TQ * j] = acc[i][j];
  } else {
    __syncthreads();  // the last slab has been consumed
    for (int q = 0; q < NG; ++q) {
      if (g == q) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int idx = (ty + TQ * i) * TL + tx + TQ * j;
            sh[idx] = q == 0 ? acc[i][j] : zadd(sh[idx], acc[i][j]);
          }
      }
      __syncthreads();
    }
    for (int idx = tid; idx < TL * TL; idx += 256) out[idx] = sh[idx];
  }
}

// The TL x TL tile (TL = 64 or 32) on the float64 matrix cores (complex128
// blocks): wave w of the block computes tile rows (TL/2) (w / 2) ..
// +TL/2 and columns (TL/2) (w % 2) .. +TL/2 as NT x NT products of 16 x 16
// (NT = TL / 32), with the real and the imaginary part in separate
// accumulators.  One v_mfma_f64_16x16x4f64 takes four rows of A and B: lane
// l holds row k0 + l / 16 and column l % 16 of a 16-column strip of each,
// loaded straight from the row-major blocks (a strip of one row is 256
// contiguous bytes); result entry r of lane l is row l / 16 + 4 r, column
// l % 16 of the 16 x 16 product (checked on the device against a host
// product).  Sixteen rows are in flight while the previous sixteen multiply.
// Partial layout as k_gram with one group.
typedef double gram_d4 __attribute__((ext_vector_type(4)));
// Row `row` of this lane's NT strips of A and B (zero beyond r1 or L).
template <int NT>
__device__ __forceinline__ void gram_fetch(const double2* __restrict__ A, const double2* __restrict__ B, long long row,
                                           long long r1, int L, int ja, int lb, double2 (&a)[NT], double2 (&b)[NT]) {
  const bool ok = row < r1;
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    a[i] = (ok && ja + 16 * i < L) ? A[row * L + ja + 16 * i] : make_double2(0.0, 0.0);
    b[i] = (ok && lb + 16 * i < L) ? B[row * L + lb + 16 * i] : make_double2(0.0, 0.0);
  }
}
template <bool CONJ, int TL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_gram_mfma(const double2* __restrict__ A, const double2* __restrict__ B,
                                                   long long rows, int L, double2* __restrict__ partial) {
  constexpr int WS = TL / 2, NT = WS / 16;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int kr = lane >> 4, cc = lane & 15;
  const int ntile = (L + TL - 1) / TL;
  const int jt = blockIdx.y / ntile, lt = blockIdx.y % ntile;
  const long long chunk = (rows + gridDim.x - 1) / gridDim.x;
  const long long r0 = blockIdx.x * chunk;
  const long long r1 = min(rows, r0 + chunk);
  const int ja = jt * TL + WS * (w >> 1) + cc, lb = lt * TL + WS * (w & 1) + cc;
  gram_d4 cr[NT][NT], ci[NT][NT];
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      cr[i][j] = gram_d4{0.0, 0.0, 0.0, 0.0};
      ci[i][j] = gram_d4{0.0, 0.0, 0.0, 0.0};
    }
  // PF groups of four rows in flight: the next 4 PF rows are loaded while
  // the current ones multiply
  constexpr int PF = 4;
  double2 a[PF][NT], b[PF][NT];
#pragma unroll
  for (int f = 0; f < PF; ++f) gram_fetch<NT>(A, B, r0 + 4 * f + kr, r1, L, ja, lb, a[f], b[f]);
  for (long long k0 = r0; k0 < r1; k0 += 4 * PF) {
    double2 ca[PF][NT], cb[PF][NT];
#pragma unroll
    for (int f = 0; f < PF; ++f)
#pragma unroll
      for (int i = 0; i < NT; ++i) {
        ca[f][i] = a[f][i];
        cb[f][i] = b[f][i];
      }
    if (k0 + 4 * PF < r1) {
#pragma unroll
      for (int f = 0; f < PF; ++f) gram_fetch<NT>(A, B, k0 + 4 * (PF + f) + kr, r1, L, ja, lb, a[f], b[f]);
    }
    // conj:  Re += ar br + ai bi, Im += ar bi - ai br
    // plain: Re += ar br - ai bi, Im += ar bi + ai br
#pragma unroll
    for (int f = 0; f < PF; ++f) {
#pragma unroll
      for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[f][i].x, cb[f][j].x, cr[i][j], 0, 0, 0);
          ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[f][i].x, cb[f][j].y, ci[i][j], 0, 0, 0);
        }
#pragma unroll
      for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          cr[i][j] =
              __builtin_amdgcn_mfma_f64_16x16x4f64(CONJ ? ca[f][i].y : -ca[f][i].y, cb[f][j].y, cr[i][j], 0, 0, 0);
          ci[i][j] =
              __builtin_amdgcn_mfma_f64_16x16x4f64(CONJ ? -ca[f][i].y : ca[f][i].y, cb[f][j].x, ci[i][j], 0, 0, 0);
        }
    }
  }
  double2* out = partial + ((long long)blockIdx.x * gridDim.y + blockIdx.y) * (TL * TL);
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int jl = WS * (w >> 1) + 16 * i + kr + 4 * r, ll = WS * (w & 1) + 16 * j + cc;
        out[jl * TL + ll] = make_double2(cr[i][j][r], ci[i][j][r]);
      }
}

// out[j][l] = sum of the parts' TL x TL tiles.  A block of 256 threads takes
// E consecutive tile entries and G = 256 / E part groups (E chosen by the
// launcher so that each thread's chain over parts stays short): group g sums
// parts g, g + G, ..., and the G group sums are added in group order (a
// fixed order: deterministic).  Entries outside the L x L result are
// skipped.
__global__ void __launch_bounds__(256) k_gram_reduce(const double2* __restrict__ partial, int nparts, int ntile,
                                                     int TL, int L, int E, double2* __restrict__ out) {
  __shared__ double2 sh[256];
  const int G = 256 / E;
  const int e = threadIdx.x % E, g = threadIdx.x / E;
  const long long id = (long long)blockIdx.x * E + e;
  const int te = TL * TL, ntile2 = ntile * ntile;
  const int tile = (int)(id / te), w = (int)(id % te);
  const int j = (tile / ntile) * TL + w / TL, l = (tile % ntile) * TL + w % TL;
  const bool live = id < (long long)ntile2 * te && j < L && l < L;
  double2 s = make_double2(0.0, 0.0);
  if (live) {
    // eight parts' loads in flight at once; the sum keeps part order
    for (int p0 = g; p0 < nparts; p0 += 8 * G) {
      double2 v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int p = p0 + u * G;
        v[u] = p < nparts ? partial[((long long)p * ntile2 + tile) * te + w] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (p0 + u * G < nparts) s = zadd(s, v[u]);
    }
  }
  sh[g * E + e] = s;
  __syncthreads();
  if (g == 0 && live) {
    for (int q = 1; q < G; ++q) s = zadd(s, sh[q * E + e]);
    out[j * L + l] = s;
  }
}

// Entries per k_gram_reduce block for `parts` partial tiles: G = 256 / E
// groups with chains of about 8 parts each, so that the reduction spreads
// over many blocks.
static int gram_reduce_width(int parts) {
  int G = 1;
  while (G < 256 && parts / G > 8) G *= 2;
  return 256 / G;
}

// Output tile of the Gram and update kernels: the smallest of 16, 32, 64
// that holds L (64 beyond).
static int block_tile(int L) { return L <= 16 ? 16 : (L <= 32 ? 32 : 64); }

template <typename V, int TL>
static void launch_gram_tile(const V* A, const V* B, long long rows, int L, bool conjA, double2* partial,
                             int nparts, double2* out, hipStream_t s) {
  // complex128 blocks with L > 32: the matrix-core kernel (one group).  On
  // the 32-column tile it measured slower than the vector kernel (that Gram
  // is bound by HBM reads, and the vector kernel's four row groups keep
  // more of them in flight).
  constexpr bool mfma = std::is_same<V, double2>::value && TL == 64;
  const int ntile = (L + TL - 1) / TL;
  dim3 grid(nparts, ntile * ntile);
  if constexpr (mfma) {
    if (conjA)
      hipLaunchKernelGGL((k_gram_mfma<true, TL>), grid, dim3(256), 0, s, A, B, rows, L, partial);
    else
      hipLaunchKernelGGL((k_gram_mfma<false, TL>), grid, dim3(256), 0, s, A, B, rows, L, partial);
  } else if (conjA) {
    hipLaunchKernelGGL((k_gram<V, true, TL>), grid, dim3(256), 0, s, A, B, rows, L, partial);
  } else {
    hipLaunchKernelGGL((k_gram<V, false, TL>), grid, dim3(256), 0, s, A, B, rows, L, partial);
  }
  const int parts = nparts;
  const int E = gram_reduce_width(parts);
  const long long entries = (long long)ntile * ntile * TL * TL;
  hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)((entries + E - 1) / E)), dim3(256), 0, s, partial, parts, ntile,
                     TL, L, E, out);
}

// The partial buffer holds nparts * 4096 * ceil(L / 64)^2 entries, which
// covers every tile: (nparts * NG) * TL^2 = nparts * 4096 for TL < 64.
template <typename V>
void launch_gram(const V* A, const V* B, long long rows, int L, bool conjA, double2* partial, int nparts,
                 double2* out, hipStream_t s) {
  switch (block_tile(L)) {
    case 16: launch_gram_tile<V, 16>(A, B, rows, L, conjA, partial, nparts, out, s); break;
    case 32: launch_gram_tile<V, 32>(A, B, rows, L, conjA, partial, nparts, out, s); break;
    default: launch_gram_tile<V, 64>(A, B, rows, L, conjA, partial, nparts, out, s); break;
  }
}
template void launch_gram<double2>(const double2*, const double2*, long long, int, bool, double2*, int, double2*,
                                   hipStream_t);
template void launch_gram<float2>(const float2*, const float2*, long long, int, bool, double2*, int, double2*,
                                  hipStream_t);

// ------------------------------------------------------------ block update
// Y = a X + (W1 + w2c W2) C + e E.  A block owns a slab of RB = 4096 / TL
// rows and a TL-column output tile (TL = 64, 32 or 16, the smallest that
// holds L, so that small L do not multiply padding); each of its 256 threads
// a 4 x 4 register tile (rows ty + RB/4 i, columns tx + TL/4 t), so one pass
// over j reads 4 W and 4 C values from LDS for 16 complex products.  j runs
// in chunks of KC (W chunk RB x KC+1 with a padded row, C chunk KC x TL), in
// increasing order, so every output is summed in the order j = 0, ..., L-1.
// The L x L coefficients, the combination and the accumulation stay in
// float64 for both vector types (gfx950 issues float64 FMAs at the float32
// rate), so a complex64 update is rounded once, at the store: with
// ill-conditioned coefficients a float accumulation would amplify rounding.
// Y may alias X (each output reads its own X entry only), not W1, W2 or E.
template <typename V, int TL>
__global__ void __launch_bounds__(256) k_update(V* Y, const V* X, double2 a, const V* W1, const V* W2,
                                                double2 w2c, const double2* __restrict__ C, const V* E,
                                                double2 e, long long rows, int L, const double2* sc) {
  using T = typename RealOf<V>::T;
  if (sc) {
    w2c = sc[0];
    e = sc[1];
  }
  constexpr int TQ = TL / 4, RB = 4096 / TL, RQ = RB / 4, KC = TL == 16 ? 8 : 16;
  __shared__ double2 Ws[RB][KC + 1];
  __shared__ double2 Cs[KC][TL];
  const int tid = threadIdx.x, tx = tid % TQ, ty = tid / TQ;
  const int col0 = blockIdx.y * TL;
  for (long long r0 = blockIdx.x * (long long)RB; r0 < rows; r0 += (long long)RB * gridDim.x) {
    double2 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[i][t] = make_double2(0.0, 0.0);
    for (int j0 = 0; j0 < L; j0 += KC) {
      __syncthreads();  // previous chunk consumed
#pragma unroll
      for (int u = 0; u < (RB * KC) / 256; ++u) {
        const int idx = tid + 256 * u, rr = idx / KC, jj = idx % KC;
        const long long row = r0 + rr;
        const int j = j0 + jj;
        double2 w = make_double2(0.0, 0.0);
        if (row < rows && j < L) {
          const V w1 = W1[row * L + j];
          w = make_double2((double)w1.x, (double)w1.y);
          if (W2) {
            const V w2 = W2[row * L + j];
            w.x += w2c.x * (double)w2.x - w2c.y * (double)w2.y;
            w.y += w2c.x * (double)w2.y + w2c.y * (double)w2.x;
          }
        }
        Ws[rr][jj] = w;
      }
      for (int idx = tid; idx < KC * TL; idx += 256) {
        const int jj = idx / TL, cc = idx % TL;
        const int j = j0 + jj, l = col0 + cc;
        Cs[jj][cc] = (j < L && l < L) ? C[(long long)j * L + l] : make_double2(0.0, 0.0);
      }
      __syncthreads();
      const int jn = min(KC, L - j0);
      for (int jj = 0; jj < jn; ++jj) {
        double2 w[4], c[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) w[i] = Ws[ty + RQ * i][jj];
#pragma unroll
        for (int t = 0; t < 4; ++t) c[t] = Cs[jj][tx + TQ * t];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            acc[i][t].x += w[i].x * c[t].x - w[i].y * c[t].y;
            acc[i][t].y += w[i].x * c[t].y + w[i].y * c[t].x;
          }
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const long long row = r0 + ty + RQ * i;
      if (row >= rows) continue;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int l = col0 + tx + TQ * t;
        if (l >= L) continue;
        const long long o = row * L + l;
        double2 y = acc[i][t];
        if (X) {
          const V x = X[o];
          y.x += a.x * (double)x.x - a.y * (double)x.y;
          y.y += a.x * (double)x.y + a.y * (double)x.x;
        }
        if (E) {
          const V ev = E[o];
          y.x += e.x * (double)ev.x - e.y * (double)ev.y;
          y.y += e.x * (double)ev.y + e.y * (double)ev.x;
        }
        V out;
        out.x = (T)y.x;
        out.y = (T)y.y;
        Y[o] = out;
      }
    }
  }
}

// The update for complex128 blocks with L <= 16 NT (NT = 1, 2 or 4) on the
// float64 matrix cores, optionally (NT <= 2) with the Gram matrix op(Y)^T Y
// of the block it writes (GRAM 1: Y^H Y, 2: Y^T Y), so that the Gram needs
// no second pass over Y.  C (L x L) is staged in LDS once per block; the
// block walks 64-row slabs, wave w taking rows 16 w .. +16 against all 16 NT
// columns.  Step st takes coefficients j = 4 st .. 4 st + 3: lane l holds
// row l % 16, coefficient 4 st + l / 16 of W1 + w2c W2 (the A operand) and
// C[4 st + l / 16][16 t + l % 16] (the B operand) of v_mfma_f64_16x16x4f64,
// real and imaginary parts in separate accumulators; sums over j run in a
// fixed order.  Every load of a slab is issued before the first is used.
// Entry r of tile t of lane
// l is Y[row kr + 4 r][column 16 t + cc] (kr = l / 16, cc = l % 16), which is
// exactly the operand layout of v_mfma_f64_16x16x4f64 over the four rows
// kr + 4 r (A[i][k] = op(Y[k][16 J + i]) from lane i + 16 k, B[k][j] =
// Y[k][16 T + j] from the same lane): the Gram tiles (J <= T) accumulate from
// the lanes' own registers.  Rows beyond `rows` hold zero.  At the end the
// four waves' tiles are summed through LDS in the fixed order
// ((w3 + w2) + w1) + w0 and written as the block's partial (TL x TL, the
// lower tiles mirrored: conjugated for GRAM 1), which k_gram_reduce adds up
// in a fixed order: repeated calls agree bit for bit.
template <int NT, int GRAM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NT == 1 ? 4 : 2))) k_update_gram(double2* Y, const double2* X, double2 a, const double2* W1,
                                                     const double2* W2, double2 w2c,
                                                     const double2* __restrict__ C, const double2* E, double2 e,
                                                     long long rows, int L, const double2* sc,
                                                     double2* __restrict__ partial) {
  constexpr int TL = 16 * NT, KS = TL / 4, NP = NT * (NT + 1) / 2, CH = NT == 4 ? 2 : 4;
  constexpr bool CONJ = GRAM == 1;
  static_assert(GRAM == 0 || NT <= 2, "the Gram accumulators of wider tiles do not fit the registers");
  __shared__ double2 Cs[TL][TL];
  if (sc) {
    w2c = sc[0];
    e = sc[1];
  }
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int kr = lane >> 4, cc = lane & 15;
  for (int idx = tid; idx < TL * TL; idx += 256) {
    const int j = idx / TL, l = idx % TL;
    Cs[j][l] = (j < L && l < L) ? C[(long long)j * L + l] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  gram_d4 gr[NP], gi[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    gr[p] = gram_d4{0.0, 0.0, 0.0, 0.0};
    gi[p] = gram_d4{0.0, 0.0, 0.0, 0.0};
  }
  const bool has_w2 = W2 != nullptr, has_x = X != nullptr, has_e = E != nullptr;
  for (long long r0 = blockIdx.x * 64LL; r0 < rows; r0 += 64LL * gridDim.x) {
    // Every load of the slab is issued before the first use, from clamped
    // (always valid) addresses; out-of-range entries are zeroed afterwards.
    const long long row = r0 + 16 * w + cc;  // this lane's W row
    const long long rowc = row < rows ? row : rows - 1;
    // the epilogue's a X + e E at entries (row kr + 4 r, column 16 t + cc),
    // loaded after the first chunk of W (in flight during its products)
    double2 xv[NT][4];
    auto entry = [&](int t, int r) -> long long {
      long long orow = r0 + 16 * w + kr + 4 * r;
      orow = orow < rows ? orow : rows - 1;
      const int l = 16 * t + cc;
      return orow * L + (l < L ? l : L - 1);
    };
    gram_d4 yr[NT], yi[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      yr[t] = gram_d4{0.0, 0.0, 0.0, 0.0};
      yi[t] = gram_d4{0.0, 0.0, 0.0, 0.0};
    }
    // W in chunks of CH coefficient steps: a chunk's loads are issued
    // together, one chunk ahead of its products
    auto load_chunk = [&](int c0, double2 (&v)[CH]) {
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int j = 4 * (c0 + q) + kr;
        v[q] = W1[rowc * L + (j < L ? j : L - 1)];
      }
      if (has_w2) {
        double2 u[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const int j = 4 * (c0 + q) + kr;
          u[q] = W2[rowc * L + (j < L ? j : L - 1)];
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          v[q].x += w2c.x * u[q].x - w2c.y * u[q].y;
          v[q].y += w2c.x * u[q].y + w2c.y * u[q].x;
        }
      }
    };
    double2 va[CH], vb[CH];
    load_chunk(0, va);
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) xv[t][r] = has_x ? X[entry(t, r)] : make_double2(0.0, 0.0);
#pragma unroll
    for (int c0 = 0; c0 < KS; c0 += CH) {
      if (c0 + CH < KS) load_chunk(c0 + CH, vb);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int j = 4 * (c0 + q) + kr;
        const double2 v = (row < rows && j < L) ? va[q] : make_double2(0.0, 0.0);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const double2 ct = Cs[j][16 * t + cc];
          yr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(v.x, ct.x, yr[t], 0, 0, 0);
          yi[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(v.x, ct.y, yi[t], 0, 0, 0);
          yr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-v.y, ct.y, yr[t], 0, 0, 0);
          yi[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(v.y, ct.x, yi[t], 0, 0, 0);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < CH; ++q) va[q] = vb[q];
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double2 x = xv[t][r];
        xv[t][r] = make_double2(a.x * x.x - a.y * x.y, a.x * x.y + a.y * x.x);
      }
    if (has_e) {
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double2 ee = E[entry(t, r)];
          xv[t][r].x += e.x * ee.x - e.y * ee.y;
          xv[t][r].y += e.x * ee.y + e.y * ee.x;
        }
    }
    double2 yv[NT][4];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long orow = r0 + 16 * w + kr + 4 * r;
        const int l = 16 * t + cc;
        const double2 y = make_double2(yr[t][r] + xv[t][r].x, yi[t][r] + xv[t][r].y);
        const bool live = orow < rows && l < L;
        if (live) Y[orow * L + l] = y;
        yv[t][r] = live ? y : make_double2(0.0, 0.0);
      }
    if constexpr (GRAM != 0) {
      // conj:  Re += ar br + ai bi, Im += ar bi - ai br
      // plain: Re += ar br - ai bi, Im += ar bi + ai br
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int p = 0;
#pragma unroll
        for (int J = 0; J < NT; ++J)
#pragma unroll
          for (int T = J; T < NT; ++T, ++p) {
            const double2 av = yv[J][r], bv = yv[T][r];
            gr[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.x, gr[p], 0, 0, 0);
            gi[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.y, gi[p], 0, 0, 0);
            gr[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(CONJ ? av.y : -av.y, bv.y, gr[p], 0, 0, 0);
            gi[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(CONJ ? -av.y : av.y, bv.x, gi[p], 0, 0, 0);
          }
      }
    }
  }
  if constexpr (GRAM != 0) {
    // ((w3 + w2) + w1) + w0 through Cs (free now), then the block's partial
    __syncthreads();
    for (int src = 3; src >= 0; --src) {
      if (w == src) {
        int p = 0;
#pragma unroll
        for (int J = 0; J < NT; ++J)
#pragma unroll
          for (int T = J; T < NT; ++T, ++p)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int j = 16 * J + kr + 4 * r, l = 16 * T + cc;
              double2 g = make_double2(gr[p][r], gi[p][r]);
              if (src < 3) g = zadd(Cs[j][l], g);
              if (src > 0) {
                Cs[j][l] = g;
              } else {
                double2* out = partial + (long long)blockIdx.x * (TL * TL);
                out[j * TL + l] = g;
                if (J < T) out[l * TL + j] = CONJ ? make_double2(g.x, -g.y) : g;
              }
            }
      }
      __syncthreads();
    }
  }
}

// Blocks of the matrix-core update (and Gram) kernel for complex128 blocks
// with L <= 32: one per 64-row slab, at most kUpdateGramBlocks (two waves
// per SIMD on every CU for the 32-column kernels).
constexpr int kUpdateGramBlocks = 512;
int update_gram_parts(long long rows) { return (int)std::min<long long>((rows + 63) / 64, kUpdateGramBlocks); }

template <int NT>
static void launch_update_gram_nt(double2* Y, const double2* X, double2 a, const double2* W1, const double2* W2,
                                  double2 w2c, const double2* C, const double2* E, double2 e, long long rows, int L,
                                  int gram, double2* partial, double2* out, hipStream_t s, const double2* sc) {
  const int g = update_gram_parts(rows);
  switch (gram) {
    case 1:
      hipLaunchKernelGGL((k_update_gram<NT, 1>), dim3(g), dim3(256), 0, s, Y, X, a, W1, W2, w2c, C, E, e, rows, L,
                         sc, partial);
      break;
    case 2:
      hipLaunchKernelGGL((k_update_gram<NT, 2>), dim3(g), dim3(256), 0, s, Y, X, a, W1, W2, w2c, C, E, e, rows, L,
                         sc, partial);
      break;
    default:
      hipLaunchKernelGGL((k_update_gram<NT, 0>), dim3(g), dim3(256), 0, s, Y, X, a, W1, W2, w2c, C, E, e, rows, L,
                         sc, partial);
      return;
  }
  constexpr int TL = 16 * NT;
  const int E_ = gram_reduce_width(g);
  hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)((TL * TL + E_ - 1) / E_)), dim3(256), 0, s, partial, g, 1, TL,
                     L, E_, out);
}

void launch_update_gram(double2* Y, const double2* X, double2 a, const double2* W1, const double2* W2, double2 w2c,
                        const double2* C, const double2* E, double2 e, long long rows, int L, int gram,
                        double2* partial, double2* out, hipStream_t s, const double2* sc) {
  if (L <= 16) {
    launch_update_gram_nt<1>(Y, X, a, W1, W2, w2c, C, E, e, rows, L, gram, partial, out, s, sc);
  } else if (L <= 32) {
    launch_update_gram_nt<2>(Y, X, a, W1, W2, w2c, C, E, e, rows, L, gram, partial, out, s, sc);
  } else {  // L <= 64, the update alone
    const int g = update_gram_parts(rows);
    hipLaunchKernelGGL((k_update_gram<4, 0>), dim3(g), dim3(256), 0, s, Y, X, a, W1, W2, w2c, C, E, e, rows, L, sc,
                       partial);
  }
}

template <typename V, int TL>
static void launch_update_tile(V* Y, const V* X, double2 a, const V* W1, const V* W2, double2 w2c,
                               const double2* C, const V* E, double2 e, long long rows, int L, hipStream_t s,
                               const double2* sc) {
  constexpr int RB = 4096 / TL;
  const long long slabs = (rows + RB - 1) / RB;
  const int g = (int)min(slabs, 4096LL);
  dim3 grid(g, (L + TL - 1) / TL);
  hipLaunchKernelGGL((k_update<V, TL>), grid, dim3(256), 0, s, Y, X, a, W1, W2, w2c, C, E, e, rows, L, sc);
}

template <typename V>
void launch_update(V* Y, const V* X, double2 a, const V* W1, const V* W2, double2 w2c, const double2* C,
                   const V* E, double2 e, long long rows, int L, hipStream_t s, const double2* sc) {
  if constexpr (std::is_same<V, double2>::value) {
    if (L <= 64) {  // the matrix-core kernel without its Gram
      launch_update_gram(Y, X, a, W1, W2, w2c, C, E, e, rows, L, 0, nullptr, nullptr, s, sc);
      return;
    }
  }
  switch (block_tile(L)) {
    case 16: launch_update_tile<V, 16>(Y, X, a, W1, W2, w2c, C, E, e, rows, L, s, sc); break;
    case 32: launch_update_tile<V, 32>(Y, X, a, W1, W2, w2c, C, E, e, rows, L, s, sc); break;
    default: launch_update_tile<V, 64>(Y, X, a, W1, W2, w2c, C, E, e, rows, L, s, sc); break;
  }
}
template void launch_update<double2>(double2*, const double2*, double2, const double2*, const double2*, double2,
                                     const double2*, const double2*, double2, long long, int, hipStream_t,
                                     const double2*);
template void launch_update<float2>(float2*, const float2*, double2, const float2*, const float2*, double2,
                                    const double2*, const float2*, double2, long long, int, hipStream_t,
                                    const double2*);

// ------------------------------------------------------------ reductions
constexpr int kDotParts = 1024;
int dot_parts() { return kDotParts; }

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x / 64); ++i) t += sh[i];
  }
  __syncthreads();
  return t;
}

template <typename V>
__global__ void __launch_bounds__(256) k_axpy_norm(V* Y, const V* X, double2 c, const V* Z, long long n,
                                                   double* partial, const double2* c_dev) {
  using T = typename RealOf<V>::T;
  if (c_dev) c = *c_dev;
  __shared__ double sh[4];
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const V x = X[i], z = Z[i];
    V y;
    y.x = x.x + (T)c.x * z.x - (T)c.y * z.y;
    y.y = x.y + (T)c.x * z.y + (T)c.y * z.x;
    Y[i] = y;
    s += (double)y.x * y.x + (double)y.y * y.y;
  }
  const double t = block_sum(s, sh);
  if (partial && threadIdx.x == 0) partial[blockIdx.x] = t;
}

template <typename V>
void launch_axpy_norm(V* Y, const V* X, double2 c, const V* Z, long long n, double* partial, int nparts,
                      hipStream_t s, const double2* c_dev) {
  hipLaunchKernelGGL((k_axpy_norm<V>), dim3(nparts), dim3(256), 0, s, Y, X, c, Z, n, partial, c_dev);
}
template void launch_axpy_norm<double2>(double2*, const double2*, double2, const double2*, long long, double*, int,
                                        hipStream_t, const double2*);
template void launch_axpy_norm<float2>(float2*, const float2*, double2, const float2*, long long, double*, int,
                                       hipStream_t, const double2*);

template <typename V>
__global__ void __launch_bounds__(256) k_dot(const V* A, const V* B, long long n, double2* pdot, double* pnrm) {
  __shared__ double sh[4];
  double sr = 0.0, si = 0.0, sn = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double2 b = d2(B[i]);
    if (pdot) {
      const double2 a = d2(A[i]);
      sr += a.x * b.x + a.y * b.y;
      si += a.x * b.y - a.y * b.x;
    }
    sn += b.x * b.x + b.y * b.y;
  }
  const double tr = block_sum(sr, sh);
  const double ti = block_sum(si, sh);
  const double tn = block_sum(sn, sh);
  if (threadIdx.x == 0) {
    if (pdot) pdot[blockIdx.x] = make_double2(tr, ti);
    if (pnrm) pnrm[blockIdx.x] = tn;
  }
}

template <typename V>
void launch_dot(const V* A, const V* B, long long n, double2* pdot, double* pnrm, int nparts, hipStream_t s) {
  hipLaunchKernelGGL((k_dot<V>), dim3(nparts), dim3(256), 0, s, A, B, n, pdot, pnrm);
}
template void launch_dot<double2>(const double2*, const double2*, long long, double2*, double*, int, hipStream_t);
template void launch_dot<float2>(const float2*, const float2*, long long, double2*, double*, int, hipStream_t);

// ------------------------------------------------------------ elementwise
template <typename V>
__global__ void k_scale_rows(V* Y, const V* X, const double2* d, long long rows, int L) {
  const long long n = rows * L;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    Y[i] = from_d2<V>(zmul(d[i / L], d2(X[i])));
  }
}
template <typename V>
void launch_scale_rows(V* Y, const V* X, const double2* d, long long rows, int L, hipStream_t s) {
  hipLaunchKernelGGL((k_scale_rows<V>), dim3(2048), dim3(256), 0, s, Y, X, d, rows, L);
}
template void launch_scale_rows<double2>(double2*, const double2*, const double2*, long long, int, hipStream_t);
template void launch_scale_rows<float2>(float2*, const float2*, const double2*, long long, int, hipStream_t);

template <typename Vd, typename Vs>
__global__ void k_convert(Vd* dst, const Vs* src, long long n, double scale) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double2 v = d2(src[i]);
    dst[i] = from_d2<Vd>(make_double2(v.x * scale, v.y * scale));
  }
}
template <typename Vd, typename Vs>
void launch_convert(Vd* dst, const Vs* src, long long n, double scale, hipStream_t s) {
  hipLaunchKernelGGL((k_convert<Vd, Vs>), dim3(2048), dim3(256), 0, s, dst, src, n, scale);
}
template void launch_convert<float2, double2>(float2*, const double2*, long long, double, hipStream_t);
template void launch_convert<double2, float2>(double2*, const float2*, long long, double, hipStream_t);
template void launch_convert<double2, double2>(double2*, const double2*, long long, double, hipStream_t);
template void launch_convert<float2, float2>(float2*, const float2*, long long, double, hipStream_t);

__global__ void k_add_c64(double2* X, const float2* D, long long n, double scale) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float2 d = D[i];
    X[i].x += scale * (double)d.x;
    X[i].y += scale * (double)d.y;
  }
}
void launch_add_c64_to_c128(double2* X, const float2* D, long long n, double scale, hipStream_t s) {
  hipLaunchKernelGGL(k_add_c64, dim3(2048), dim3(256), 0, s, X, D, n, scale);
}

__global__ void k_fill_int(int* p, long long n, int v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_scatter_inv(int* inv, const long long* addr, long long N) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N;
       i += (long long)gridDim.x * blockDim.x)
    inv[addr[i]] = (int)i;
}
void launch_fill_inv(int* inv, long long ncell, const long long* addr, long long N, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_int, dim3(1024), dim3(256), 0, s, inv, ncell, -1);
  hipLaunchKernelGGL(k_scatter_inv, dim3(1024), dim3(256), 0, s, inv, addr, N);
}

// -------------------------------------------------------------- physics
// B[3m+c][l] = exp(i k u_l . r_m) e0[l][c]
__global__ void k_rhs(double2* B, const double* pos, const double* u, const double2* e0, double k, long long N,
                      int L) {
  const long long n = N * L;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < n;
       id += (long long)gridDim.x * blockDim.x) {
    const long long m = id / L;
    const int l = id % L;
    const double ph = k * (u[3 * l] * pos[3 * m] + u[3 * l + 1] * pos[3 * m + 1] + u[3 * l + 2] * pos[3 * m + 2]);
    double sn, cs;
    sincos(ph, &sn, &cs);
    const double2 e = make_double2(cs, sn);
    for (int c = 0; c < 3; ++c) B[(3 * m + c) * L + l] = zmul(e, e0[3 * l + c]);
  }
}
void launch_rhs(double2* B, const double* pos, const double* u, const double2* e0, double k, long long N, int L,
                hipStream_t s) {
  hipLaunchKernelGGL(k_rhs, dim3(2048), dim3(256), 0, s, B, pos, u, e0, k, N, L);
}

// Per orientation l and per workgroup: sum over its dipoles of
//  [0] P.th_fw conj(e_fw)   [1] P.ph_fw (-i) conj(e_fw)
//  [2] P.th_bk conj(e_bk)   [3] P.ph_bk (-i) conj(e_bk)
//  [4] (sum_c Im(P_c conj(P_c f_c)), sum_c Im(P_c conj(e_fw e0_c)))
// geo per l: u_fw(3) th_fw(3) ph_fw(3) u_bk(3) th_bk(3) ph_bk(3) e0 (3 complex = 6)
constexpr int kObsParts = 256;
int obs_parts() { return kObsParts; }

__global__ void __launch_bounds__(256) k_observables(const double2* __restrict__ X, const double* __restrict__ pos,
                                                     const double* __restrict__ geo,
                                                     const double2* __restrict__ epsfac, double k, long long N,
                                                     int L, double2* __restrict__ partial) {
  __shared__ double2 red[4][64][5];
  const int tl = threadIdx.x % 64, tm = threadIdx.x / 64;
  for (int l0 = 0; l0 < L; l0 += 64) {
    const int l = l0 + tl;
    double2 acc[5];
    for (int t = 0; t < 5; ++t) acc[t] = make_double2(0.0, 0.0);
    if (l < L) {
      const double* g = geo + 24 * l;
      for (long long m = (long long)blockIdx.x * 4 + tm; m < N; m += 4LL * gridDim.x) {
        const double rx = pos[3 * m], ry = pos[3 * m + 1], rz = pos[3 * m + 2];
        double2 P[3];
        for (int c = 0; c < 3; ++c) P[c] = X[(3 * m + c) * L + l];
        double sn, cs;
        sincos(k * (g[0] * rx + g[1] * ry + g[2] * rz), &sn, &cs);
        const double2 efw = make_double2(cs, sn);   // e_fw (= e_inc phase)
        const double2 cef = make_double2(cs, -sn);  // conj(e_fw)
        sincos(k * (g[9] * rx + g[10] * ry + g[11] * rz), &sn, &cs);
        const double2 ceb = make_double2(cs, -sn);  // conj(e_bk)
        double2 pt = make_double2(0, 0), pp = make_double2(0, 0), bt = make_double2(0, 0), bp = make_double2(0, 0);
        for (int c = 0; c < 3; ++c) {
          pt.x += P[c].x * g[3 + c];
          pt.y += P[c].y * g[3 + c];
          pp.x += P[c].x * g[6 + c];
          pp.y += P[c].y * g[6 + c];
          bt.x += P[c].x * g[12 + c];
          bt.y += P[c].y * g[12 + c];
          bp.x += P[c].x * g[15 + c];
          bp.y += P[c].y * g[15 + c];
        }
        acc[0] = zadd(acc[0], zmul(pt, cef));
        acc[1] = zadd(acc[1], zmul(zmul(pp, make_double2(0.0, -1.0)), cef));
        acc[2] = zadd(acc[2], zmul(bt, ceb));
        acc[3] = zadd(acc[3], zmul(zmul(bp, make_double2(0.0, -1.0)), ceb));
        for (int c = 0; c < 3; ++c) {
          const double2 E = zmul(P[c], epsfac[c]);
          acc[4].x += P[c].y * E.x - P[c].x * E.y;  // Im(P conj(E))
          const double2 Ei = zmul(efw, make_double2(g[18 + 2 * c], g[19 + 2 * c]));
          acc[4].y += P[c].y * Ei.x - P[c].x * Ei.y;  // Im(P conj(E_inc))
        }
      }
    }
    for (int t = 0; t < 5; ++t) red[tm][tl][t] = acc[t];
    __syncthreads();
    if (tm == 0 && l < L) {
      for (int t = 0; t < 5; ++t) {
        double2 s = red[0][tl][t];
        for (int w = 1; w < 4; ++w) s = zadd(s, red[w][tl][t]);
        partial[((long long)blockIdx.x * L + l) * 5 + t] = s;
      }
    }
    __syncthreads();
  }
}

void launch_observables(const double2* X, const double* pos, const double* geo, const double2* epsfac, double k,
                        long long N, int L, double2* partial, int nparts, hipStream_t s) {
  hipLaunchKernelGGL(k_observables, dim3(nparts), dim3(256), 0, s, X, pos, geo, epsfac, k, N, L, partial);
}

// out[l][t] = sum over the workgroup partials in workgroup order (the
// order the host-side sum of bdda_observables uses, so results are equal).
__global__ void __launch_bounds__(256) k_obs_reduce(const double2* __restrict__ partial, int nparts, int L,
                                                    double2* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // i = l * 5 + t
  if (i >= L * 5) return;
  double sx = 0.0, sy = 0.0;
  for (int w = 0; w < nparts; ++w) {
    const double2 v = partial[(long long)w * L * 5 + i];
    sx += v.x;
    sy += v.y;
  }
  out[i] = make_double2(sx, sy);
}

void launch_obs_reduce(const double2* partial, int nparts, int L, double2* out, hipStream_t s) {
  hipLaunchKernelGGL(k_obs_reduce, dim3((L * 5 + 255) / 256), dim3(256), 0, s, partial, nparts, L, out);
}

// ---------------------------------------------------------------- voxeliser
// dist[i] = sqrt(min_j ((dx*dx + dy*dy) + dz*dz)), evaluated without
// contraction so that every squared distance is the KD-tree's bit for bit
// (the minimum is order-free and sqrt is correctly rounded).
#pragma clang fp contract(off)
__global__ void __launch_bounds__(256) k_voxel_dist(const double* __restrict__ grid, long long n,
                                                    const double* __restrict__ surf, long long P,
                                                    double* __restrict__ dist) {
  __shared__ double sp[3][1024];
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double gx = 0, gy = 0, gz = 0;
  if (i < n) {
    gx = grid[3 * i];
    gy = grid[3 * i + 1];
    gz = grid[3 * i + 2];
  }
  double best = 1.0e300;
  for (long long j0 = 0; j0 < P; j0 += 1024) {
    const int cnt = (int)min(1024LL, P - j0);
    __syncthreads();
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
      sp[0][t] = surf[3 * (j0 + t)];
      sp[1][t] = surf[3 * (j0 + t) + 1];
      sp[2][t] = surf[3 * (j0 + t) + 2];
    }
    __syncthreads();
    for (int t = 0; t < cnt; ++t) {
      const double dx = gx - sp[0][t];
      const double dy = gy - sp[1][t];
      const double dz = gz - sp[2][t];
      double r2 = dx * dx;
      r2 = r2 + dy * dy;
      r2 = r2 + dz * dz;
      best = r2 < best ? r2 : best;
    }
  }
  if (i < n) dist[i] = sqrt(best);
}
#pragma clang fp contract(on)

void launch_voxel_dist(const double* grid, long long n, const double* surf, long long P, double* dist,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_voxel_dist, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, grid, n, surf, P, dist);
}

// inside along an axis <=> a near voxel lies strictly before and strictly
// after it on that line (first_near < i < last_near); AND over the 3 axes.
__global__ void k_voxel_inside(const double* __restrict__ dist, double thresh, int nx, int ny, int nz,
                               unsigned char* __restrict__ inside) {
  const long long lines_z = (long long)nx * ny, lines_y = (long long)nx * nz, lines_x = (long long)ny * nz;
  const long long total = lines_z + lines_y + lines_x;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    long long base, stride;
    int len;
    if (id < lines_z) {
      base = id * nz;
      stride = 1;
      len = nz;
    } else if (id < lines_z + lines_y) {
      const long long t = id - lines_z;
      const long long ix = t / nz, iz = t % nz;
      base = ix * (long long)ny * nz + iz;
      stride = nz;
      len = ny;
    } else {
      const long long t = id - lines_z - lines_y;
      base = t;
      stride = (long long)ny * nz;
      len = nx;
    }
    int first = len, last = -1;
    for (int i = 0; i < len; ++i)
      if (dist[base + i * stride] < thresh) {
        if (first == len) first = i;
        last = i;
      }
    for (int i = 0; i < len; ++i) {
      if (!(first < i && i < last)) inside[base + i * stride] = 0;
    }
  }
}

__global__ void k_fill_u8(unsigned char* p, long long n, unsigned char v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

void launch_voxel_inside(const double* dist, double thresh, int nx, int ny, int nz, unsigned char* inside,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u8, dim3(1024), dim3(256), 0, s, inside, (long long)nx * ny * nz, (unsigned char)1);
  hipLaunchKernelGGL(k_voxel_inside, dim3(1024), dim3(256), 0, s, dist, thresh, nx, ny, nz, inside);
}

}  // namespace bdda
