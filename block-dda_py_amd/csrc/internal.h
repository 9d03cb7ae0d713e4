// Internal declarations shared by the kernel and host translation units of
// libbdda (MI355X / gfx950).  Not part of the public C ABI (include/bdda.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace bdda {

// ----------------------------------------------------------------- errors
void set_error(const std::string& msg);
#define BDDA_HIP(call)                                                        \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::bdda::set_error(std::string(#call) + ": " + hipGetErrorString(e_));   \
      return 2;                                                               \
    }                                                                         \
  } while (0)

// ------------------------------------------------------------ FFT plans
constexpr int kMaxRadix = 12;  // passes per axis

// One axis of the doubled (convolution) grid.
struct AxisPlan {
  int n = 0;        // cuboid cells on this axis (N)
  int m = 0;        // FFT length (>= 2N-1, 5-smooth)
  int h = 0;        // m/2 + 1: stored half-spectrum of the Green tensor
  int nrad = 0;
  int rad[kMaxRadix] = {0};
};

// Radix sequence for a 5-smooth length (radices 8, 4, 2, 3, 5).
bool factor_length(int m, AxisPlan& ax);
int smallest_5smooth_at_least(int v);
// Convolution FFT length for n cells on axis a (0 = x): 2n when 5-smooth,
// else the smallest 5-smooth length >= 2n - 1, except on x, where it is the
// smallest register-kernel length >= 2n - 1 with an even first radix (what
// the dedicated x-pass kernel k_xconv needs) when one exists.
int conv_length(int n, int axis);

// Launch-time description of one batched line-FFT pass.  Every launch moves
// `npts` points (lines) x Q columns of length-m transforms through LDS.
enum PassMode : int {
  PASS_PLAIN = 0,       // global -> FFT -> global (strided lines)
  PASS_SCATTER_Z = 1,   // occupied block (3N,L) -> z FFT -> T1
  PASS_CONV_X = 2,      // x FFT -> 3x3 Green multiply -> inverse x FFT
  PASS_GATHER_Z = 3,    // T1 -> inverse z FFT -> diag + Jacobi epilogue -> (3N,L)
  PASS_CUBOID_Z = 4,    // cuboid (3n,L) -> z FFT -> T1
  PASS_CROP_Z = 5,      // T1 -> inverse z FFT -> cuboid (3n,L) (fft_mvp)
  PASS_PLAIN_INV = 6,   // PASS_PLAIN with the inverse transform (register kernels)
  PASS_SCATTER_Z64 = 7, // PASS_SCATTER_Z reading a complex128 block (complex64 transform)
  PASS_GATHER_Z64 = 8,  // PASS_GATHER_Z with a complex128 epilogue and output (complex64 transform)
};

struct LinePass {
  int mode;
  int inverse;
  int m, nin, nout;
  int nrad;
  int rad[kMaxRadix];
  const void* tw;       // twiddle table of length m (forward sign), V-typed
  int Q;                // columns per point (3*Lb, or 6 for the Green tensor)
  int npts;             // number of points (lines)
  int ptb;              // point index = a*ptb + b
  long long sa, sb;     // point offsets (complex units)  -- source
  long long se;         // element stride (complex units)  -- source
  long long da, db, de; // same for destination
  long long lo_src, lo_dst;  // column-tile strides (complex units)
  int P;                // points per workgroup
  int order_pe;         // flat tile order: 1 -> (q, p, e), 0 -> (q, e, p)
  const void* src;
  void* dst;
  // scatter / gather (solver layout)
  const int* inv;       // cell -> occupied index or -1
  // occupied blocks: the z lines (ix * Ny + iy) holding an occupied cell.
  // The z passes run only those (point pt -> line zlines[pt]); the y inverse
  // pass writes T1 only on those (lmask, one byte per line), so T1 keeps the
  // zeros its other lines were cleared to.  Null: all.
  const int* zlines;
  const unsigned char* lmask;
  int nyz_y, nz;        // cuboid Ny, Nz (z-line points are (ix, iy))
  int L, Lb;            // total columns, column tile
  int col0;             // first column of this launch's tiles (scatter / gather blocks)
  int xsplit;           // x pass kernel: 0 the process-wide rule (set_fft_path), 1 split-spectrum
                        // where eligible, 2 one-pass, 3 split only where R2 = R1 / 2
  int nlo;              // number of column tiles (grid.y)
  const void* S;        // occupied block (3N, L) (scatter source / diag term)
  void* AS;             // occupied output (3N, L)
  const void* diag;     // (3N) complex
  const void* jpre;     // (3N) complex
  double scale;         // 1/M for the inverse transform
  // Green multiply
  const void* green;    // half spectrum, [(ay*hz + az)*hx + ax][6]
  int hx, hy, hz;
  int mx, my, mz;
};

template <typename V>
void launch_line_pass(const LinePass& lp, hipStream_t s);
// FFT kernel selection (tests compare the paths): 0 register kernels where
// a length has one (the dedicated x-pass kernels on the compiled tiles: the
// split-spectrum one where R2 = R1 / 2, the one-pass one elsewhere), 1 the
// runtime-radix kernel only, 2 register kernels without the dedicated x pass,
// 3 the split x pass on every eligible length, 4 the one-pass x pass
// everywhere.
void set_fft_path(int path);
template <typename V>
void launch_green_build(V* out, const AxisPlan* ax, double lf, double k, hipStream_t s);
template <typename V>
void launch_green_halve(const double2* full, V* half, const AxisPlan* ax, hipStream_t s);
void launch_green_sym9(const double2* g6, double2* g9, long long M, hipStream_t s);
void launch_green_expand(const double2* half, double2* full, const AxisPlan* ax, const int* n2,
                         hipStream_t s);

// --------------------------------------------------------- block kernels
// Gram: G[j][l] = sum_i op(A[i][j]) * B[i][l]   (op = conj or identity), FP64 out.
template <typename V>
void launch_gram(const V* A, const V* B, long long rows, int L, bool conjA, double2* partial,
                 int nparts, double2* out, hipStream_t s);
int gram_parts(int L, long long rows);
// Y = a*X + sum_j (W1 + w2c*W2)[i][j] C[j][l] + e*E ;  X, W2, E nullable.
template <typename V>
void launch_update(V* Y, const V* X, double2 a, const V* W1, const V* W2, double2 w2c,
                   const double2* C, const V* E, double2 e, long long rows, int L, hipStream_t s,
                   const double2* sc = nullptr);  // sc (device, nullable): w2c = sc[0], e = sc[1]
// The same update for complex128 blocks with L <= 32 on the float64 matrix
// cores; gram = 1 (Y^H Y) or 2 (Y^T Y) also writes the Gram matrix of the new
// block Y into out (L x L, device), through update_gram_parts(rows) partial
// TL x TL tiles in `partial`; gram = 0 is the update alone.
void launch_update_gram(double2* Y, const double2* X, double2 a, const double2* W1, const double2* W2, double2 w2c,
                        const double2* C, const double2* E, double2 e, long long rows, int L, int gram,
                        double2* partial, double2* out, hipStream_t s, const double2* sc = nullptr);
int update_gram_parts(long long rows);
constexpr int kUpdateGramMaxL = 32;
// Y = X + c*Z, and partial |Y|^2 sums (nullable).
template <typename V>
void launch_axpy_norm(V* Y, const V* X, double2 c, const V* Z, long long n, double* partial,
                      int nparts, hipStream_t s, const double2* c_dev = nullptr);  // c = *c_dev if given
// partial sums of tr(A^H B) and |B|^2 (each nullable) -> double2/double partials
template <typename V>
void launch_dot(const V* A, const V* B, long long n, double2* part_dot, double* part_nrm,
                int nparts, hipStream_t s);
int dot_parts();
template <typename V>
void launch_scale_rows(V* Y, const V* X, const double2* d, long long rows, int L, hipStream_t s);
template <typename Vd, typename Vs>
void launch_convert(Vd* dst, const Vs* src, long long n, double scale, hipStream_t s);
void launch_add_c64_to_c128(double2* X, const float2* D, long long n, double scale, hipStream_t s);
void launch_fill_inv(int* inv, long long ncell, const long long* addr, long long N, hipStream_t s);

// learned warm start (float32 storage, bf16 MFMA)
struct GemmArgs {
  int batch, M, N, K;
  float alpha, beta;
  const float* A;
  long long sAb, sAi, sAk;
  const float* B;
  long long sBb, sBk, sBj;
  float* C;
  long long sCb, sCi, sCj;
  const float* bias;  // per row i of C, added after the product (nullable)
  int act;            // 1: GELU of the result (after the bias), 0: none
};
void launch_gemm_bf16(const GemmArgs& g, hipStream_t s);
void launch_gelu(const float* x, float* y, long long n, hipStream_t s);
void launch_gelu_bwd(const float* x, const float* dy, float* dx, long long n, hipStream_t s);
void launch_bias_add(float* y, const float* b, long long B, int C, long long S, hipStream_t s);
void launch_fno_ycombine(const float* R, float* xr, float* xi, long long nq, int K, hipStream_t s);
void launch_bias_grad(const float* dy, float* db, long long B, int C, long long S, hipStream_t s);
void launch_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
                 float c1, float c2, hipStream_t s);

void launch_fno_inputs(float* out, const unsigned char* occ, const int* sample_inst, const float* desc, int G,
                       long long nsamp, hipStream_t s);
void launch_cube_exchange(float* cube, double2* X, const long long* cells, const double2* alpha, long long N, int L,
                          int G, int to_cube, hipStream_t s);
void launch_masked_l2(const float* pred, const float* target, const unsigned char* occ, const int* sample_inst,
                      int C, long long S, float* loss, float* grad, int nsamp, hipStream_t s);

// physics
void launch_rhs(double2* B, const double* pos, const double* u, const double2* e0, double k,
                long long N, int L, hipStream_t s);
void launch_observables(const double2* X, const double* pos, const double* geo, const double2* epsfac,
                        double k, long long N, int L, double2* partial, int nparts, hipStream_t s);
int obs_parts();

// Largest block width the solver entry points take.  The algorithms have no
// limit; this one bounds the L x L host algebra and the Gram workspace.
constexpr int kMaxL = 4096;

// --------------------------------------------- device L x L algebra (L <= 64)
constexpr int kDlaMax = 64;
enum { DLA_SINGULAR = 1, DLA_NOT_PD = 2, DLA_SHIFTED = 4 };
// X = A^-1 op(B) (n x n A, n x m B, op: negation when neg_b); Xneg = -X,
// LU / piv (for dla_lu_resolve) nullable; DLA_SINGULAR in *status, and zero
// outputs, on an exactly singular A or when *status already holds it.
void dla_lu_solve(const double2* A, const double2* B, int n, int m, bool neg_b, double2* X, double2* Xneg,
                  double2* LU, int* piv, int* status, hipStream_t s);
void dla_lu_resolve(const double2* LU, const int* piv, const double2* B, int n, int m, bool neg_b, double2* X,
                    const int* status, hipStream_t s);
// R^H R = G (+ shift I if the plain factorisation fails and shift is set);
// R upper, Ri = R^-1.
void dla_chol_inv(const double2* G, int n, double rows, bool shift, double2* R, double2* Ri, int* status,
                  hipStream_t s);
void dla_matmul(const double2* A, bool ta, const double2* B, double2* C, int n, hipStream_t s);
void dla_frob(const double2* A, int n, double scale, double* out, hipStream_t s);
void dla_xi(const double2* pdot, const double* pnrm, int np, const int* status, double2* sc, hipStream_t s);
void dla_norm_err(const double* pnrm, int np, double scale, double* out, hipStream_t s);
void launch_obs_reduce(const double2* partial, int nparts, int L, double2* out, hipStream_t s);
void launch_voxel_dist(const double* grid, long long n, const double* surf, long long P, double* dist,
                       hipStream_t s);
void launch_voxel_inside(const double* dist, double thresh, int nx, int ny, int nz, unsigned char* inside,
                         hipStream_t s);

}  // namespace bdda
