// Kernels of the learned warm start (3-D Fourier neural operator) on gfx950.
//
// Every linear stage of the network -- pointwise channel mixing, the
// truncated 3-D DFTs (as products with fixed DFT matrices), the per-mode
// spectral weights, and all their transposes in the backward pass -- is one
// batched, arbitrarily strided GEMM C = alpha * op(A) op(B) + beta * C with
// float32 storage, bfloat16 MFMA inputs (round to nearest even) and float32
// accumulation: v_mfma_f32_16x16x16_bf16 on 64 x 64 output tiles, 4 waves of
// 2 x 2 MFMA tiles each, K staged through LDS in steps of 16.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace bdda {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

template <int TILE = 64, int KS = 16>
__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
  __shared__ unsigned short As[TILE][KS + 4];
  __shared__ unsigned short Bs[KS][TILE + 4];
  const long long bz = blockIdx.z;
  const float* A = g.A + bz * g.sAb;
  const float* B = g.B + bz * g.sBb;
  float* C = g.C + bz * g.sCb;
  const int i0 = blockIdx.y * TILE, j0 = blockIdx.x * TILE;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += KS) {
#pragma unroll
    for (int t = 0; t < (TILE * KS) / 256; ++t) {
      const int idx = tid + t * 256;
      const int r = idx / KS, kk = idx % KS;
      const int gi = i0 + r, gk = k0 + kk;
      const float v = (gi < g.M && gk < g.K) ? A[gi * g.sAi + (long long)gk * g.sAk] : 0.f;
      As[r][kk] = f2bf(v);
    }
#pragma unroll
    for (int t = 0; t < (TILE * KS) / 256; ++t) {
      const int idx = tid + t * 256;
      const int kk = idx / TILE, c = idx % TILE;
      const int gk = k0 + kk, gj = j0 + c;
      const float v = (gk < g.K && gj < g.N) ? B[(long long)gk * g.sBk + gj * g.sBj] : 0.f;
      Bs[kk][c] = f2bf(v);
    }
    __syncthreads();
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
      bf16x4 af;
      const int row = wr + ti * 16 + (lane & 15);
      const int kb = 4 * (lane >> 4);
#pragma unroll
      for (int j = 0; j < 4; ++j) af[j] = (short)As[row][kb + j];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) {
        bf16x4 bfr;
        const int col = wc + tj * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) bfr[j] = (short)Bs[kb + j][col];
        acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bfr, acc[ti][tj], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = i0 + wr + ti * 16 + 4 * (lane >> 4) + j;
        const int c = j0 + wc + tj * 16 + (lane & 15);
        if (r < g.M && c < g.N) {
          float* o = C + r * g.sCi + (long long)c * g.sCj;
          const float prev = g.beta != 0.f ? *o : 0.f;
          float val = g.alpha * acc[ti][tj][j] + g.beta * prev;
          if (g.bias) val += g.bias[r];
          if (g.act) val = 0.5f * val * (1.f + erff(val * 0.70710678118654752f));
          *o = val;
        }
      }
}

void launch_gemm_bf16(const GemmArgs& g, hipStream_t s) {
  dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, g.batch);
  hipLaunchKernelGGL((k_gemm_bf16<64, 16>), grid, dim3(256), 0, s, g);
}

// ------------------------------------------------------------ elementwise
// exact GELU 0.5 x (1 + erf(x / sqrt 2)) and its derivative
__global__ void k_gelu(const float* x, float* y, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float v = x[i];
    y[i] = 0.5f * v * (1.f + erff(v * 0.70710678118654752f));
  }
}
__global__ void k_gelu_bwd(const float* x, const float* dy, float* dx, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float v = x[i];
    const float cdf = 0.5f * (1.f + erff(v * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * v * v);
    dx[i] = dy[i] * (cdf + v * pdf);
  }
}
// y[b][c][s] += bias[c]   over a (B, C, S) tensor
__global__ void k_bias_add(float* y, const float* bias, long long B, int C, long long S) {
  const long long n = B * C * S;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] += bias[(i / S) % C];
}
// db[c] = sum_{b,s} dy[b][c][s]; one workgroup per channel
__global__ void __launch_bounds__(256) k_bias_grad(const float* dy, float* db, long long B, int C, long long S) {
  __shared__ float sh[256];
  const int c = blockIdx.x;
  float acc = 0.f;
  for (long long b = 0; b < B; ++b)
    for (long long s = threadIdx.x; s < S; s += blockDim.x) acc += dy[(b * C + c) * S + s];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = sh[0];
}
// Adam step with bias correction
__global__ void k_adam(float* p, const float* gr, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float gi = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// ------------------------------------------------------------ data layout
// Network samples are (instance, orientation) pairs on a G^3 cube; cell
// s = (ix*G + iy)*G + iz.  Input channels (14):
//   0 occupancy, 1..6 chi_x, chi_y, chi_z (re, im), 7..12 E_inc x, y, z
//   (re, im) inside the particle, 13 k * lattice spacing.
// desc (per sample, 20 floats): chi (6), u (3), e0 (6), k*lf, origin (3), pad
__global__ void k_fno_inputs(float* out, const unsigned char* occ, const int* sample_inst, const float* desc,
                             float lf_over_k_unused, int G, long long nsamp) {
  const long long S = (long long)G * G * G;
  const long long n = nsamp * S;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < n;
       id += (long long)gridDim.x * blockDim.x) {
    const long long b = id / S, s = id % S;
    const float* d = desc + b * 20;
    const int inst = sample_inst[b];
    const float o = occ[(long long)inst * S + s] ? 1.f : 0.f;
    const int iz = s % G, iy = (s / G) % G, ix = s / ((long long)G * G);
    const float kl = d[15];
    // phase k u.r with r = origin + index * lf  (lf folded into k*lf)
    const float ph = kl * (d[6] * (d[16] + ix) + d[7] * (d[17] + iy) + d[8] * (d[18] + iz));
    float sn, cs;
    sincosf(ph, &sn, &cs);
    float* o_ = out + b * 14 * S + s;
    o_[0] = o;
    for (int c = 0; c < 6; ++c) o_[(1 + c) * S] = o * d[c];
    for (int c = 0; c < 3; ++c) {
      const float er = d[9 + 2 * c], ei = d[10 + 2 * c];
      o_[(7 + 2 * c) * S] = o * (cs * er - sn * ei);
      o_[(8 + 2 * c) * S] = o * (cs * ei + sn * er);
    }
    o_[13 * S] = kl;
  }
}

// cube[b][2c+{0,1}][cell] <-> X[3m+c][l] / alpha_c   (b = inst_base + l)
// to_cube: cube = X / alpha (zero elsewhere is left to the caller's fill);
// from_cube: X = alpha * cube.
__global__ void k_cube_exchange(float* cube, double2* X, const long long* cells, const double2* alpha, long long N,
                                int L, int G, int to_cube) {
  const long long S = (long long)G * G * G;
  const long long n = N * L;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < n;
       id += (long long)gridDim.x * blockDim.x) {
    const long long m = id / L;
    const int l = id % L;
    const long long cell = cells[m];
    for (int c = 0; c < 3; ++c) {
      const double2 a = alpha[c];
      float* cr = cube + ((long long)l * 6 + 2 * c) * S + cell;
      float* ci = cr + S;
      double2* x = X + (3 * m + c) * L + l;
      if (to_cube) {
        // x / a
        const double den = a.x * a.x + a.y * a.y;
        *cr = (float)((x->x * a.x + x->y * a.y) / den);
        *ci = (float)((x->y * a.x - x->x * a.y) / den);
      } else {
        const double pr = *cr, pi = *ci;
        *x = make_double2(a.x * pr - a.y * pi, a.x * pi + a.y * pr);
      }
    }
  }
}

// Masked relative L2 per sample: loss_b = |mask (p - t)|^2 / |mask t|^2 and
// grad = 2 mask (p - t) / |mask t|^2 / nsamp (mean over samples).  One
// workgroup per sample; pred/target (nsamp, C, S), mask (inst, S).
__global__ void __launch_bounds__(256) k_masked_l2(const float* pred, const float* target,
                                                   const unsigned char* occ, const int* sample_inst, int C,
                                                   long long S, float* loss, float* grad, int nsamp) {
  __shared__ float sh[2][256];
  const int b = blockIdx.x;
  const unsigned char* m = occ + (long long)sample_inst[b] * S;
  const float* p = pred + (long long)b * C * S;
  const float* t = target + (long long)b * C * S;
  float num = 0.f, den = 0.f;
  for (long long i = threadIdx.x; i < C * S; i += blockDim.x) {
    if (!m[i % S]) continue;
    const float d = p[i] - t[i];
    num += d * d;
    den += t[i] * t[i];
  }
  sh[0][threadIdx.x] = num;
  sh[1][threadIdx.x] = den;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  const float tn = sh[1][0] > 0.f ? sh[1][0] : 1.f;
  if (threadIdx.x == 0) loss[b] = sh[0][0] / tn;
  if (grad) {
    float* g = grad + (long long)b * C * S;
    const float sc = 2.f / tn / (float)nsamp;
    for (long long i = threadIdx.x; i < C * S; i += blockDim.x) g[i] = m[i % S] ? sc * (p[i] - t[i]) : 0.f;
  }
}

void launch_fno_inputs(float* out, const unsigned char* occ, const int* sample_inst, const float* desc, int G,
                       long long nsamp, hipStream_t s) {
  hipLaunchKernelGGL(k_fno_inputs, dim3(2048), dim3(256), 0, s, out, occ, sample_inst, desc, 0.f, G, nsamp);
}
void launch_cube_exchange(float* cube, double2* X, const long long* cells, const double2* alpha, long long N, int L,
                          int G, int to_cube, hipStream_t s) {
  hipLaunchKernelGGL(k_cube_exchange, dim3(2048), dim3(256), 0, s, cube, X, cells, alpha, N, L, G, to_cube);
}
void launch_masked_l2(const float* pred, const float* target, const unsigned char* occ, const int* sample_inst,
                      int C, long long S, float* loss, float* grad, int nsamp, hipStream_t s) {
  hipLaunchKernelGGL(k_masked_l2, dim3(nsamp), dim3(256), 0, s, pred, target, occ, sample_inst, C, S, loss, grad,
                     nsamp);
}

// The y DFT of the spectral layer as one wide product gives, per
// (sample, channel, x) slab q, R[r][q][c]: rows r < K hold F_re times the
// slab's z spectrum, rows K + r F_im times it; columns c < K its real parts,
// K + c its imaginary parts.  The complex product is then
// X_re[q][r][c] = R[r][q][c] - R[K + r][q][K + c],
// X_im[q][r][c] = R[r][q][K + c] + R[K + r][q][c].
__global__ void k_fno_ycombine(const float* __restrict__ R, float* __restrict__ xr, float* __restrict__ xi,
                               long long nq, int K) {
  const long long kk = (long long)K * K, n = nq * kk, rs = nq * 2 * K;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long q = i / kk;
    const int r = (int)(i % kk) / K, c = (int)(i % K);
    const float* a = R + r * rs + q * 2 * K;
    const float* b = a + K * rs;
    xr[i] = a[c] - b[K + c];
    xi[i] = a[K + c] + b[c];
  }
}
void launch_fno_ycombine(const float* R, float* xr, float* xi, long long nq, int K, hipStream_t s) {
  hipLaunchKernelGGL(k_fno_ycombine, dim3(4096), dim3(256), 0, s, R, xr, xi, nq, K);
}

void launch_gelu(const float* x, float* y, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_gelu, dim3(2048), dim3(256), 0, s, x, y, n);
}
void launch_gelu_bwd(const float* x, const float* dy, float* dx, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_gelu_bwd, dim3(2048), dim3(256), 0, s, x, dy, dx, n);
}
void launch_bias_add(float* y, const float* b, long long B, int C, long long S, hipStream_t s) {
  hipLaunchKernelGGL(k_bias_add, dim3(2048), dim3(256), 0, s, y, b, B, C, S);
}
void launch_bias_grad(const float* dy, float* db, long long B, int C, long long S, hipStream_t s) {
  hipLaunchKernelGGL(k_bias_grad, dim3(C), dim3(256), 0, s, dy, db, B, C, S);
}
void launch_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
                 float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3(2048), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
}

}  // namespace bdda
