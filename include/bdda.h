/*
 * bdda.h — C ABI of libbdda, the MI355X (gfx950) block-DDA solver.
 *
 * Drop-in boundary for the reference block-DDA_Py v0.3.0 solve path
 * (mvp_fft -> bl_krylov -> bl_dda.scatterer).  Every entry point names the
 * reference interface it replaces.  Plain pointers and sizes only:
 *
 *   *_dev   device pointers (e.g. torch CUDA tensors' data_ptr())
 *   *_host  host pointers
 *   complex values are interleaved (re, im) doubles ("c128") in the
 *   reference's layouts: Krylov blocks (3N, L) row-major with row = 3m + c,
 *   cuboid blocks (3n, L) with row = 3*(ix*Ny*Nz + iy*Nz + iz) + c.
 *
 * Every function returns BDDA_OK (0) or an error code; bdda_last_error()
 * gives the message of the last failure in the calling thread.  A plan is
 * not re-entrant: calls on one plan must not overlap.
 */
#ifndef BDDA_H_
#define BDDA_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bdda_plan bdda_plan;

enum {
  BDDA_OK = 0,
  BDDA_E_ARG = 1,      /* invalid argument (the reference raises ValueError)      */
  BDDA_E_HIP = 2,      /* HIP runtime failure                                       */
  BDDA_E_SINGULAR = 3, /* singular L x L system (numpy.linalg.LinAlgError)          */
  BDDA_E_OOM = 4,      /* device allocation failed                                  */
  BDDA_E_STATE = 5     /* call order (e.g. solve before green_init / set_target)    */
};

enum { BDDA_SOLVER_COCG_RQ = 0, BDDA_SOLVER_BICGSTAB = 1 };
enum { BDDA_PREC_FP64 = 0, BDDA_PREC_MIXED = 1 };

/* Message of the last failed call in this thread ("" if none). */
const char* bdda_last_error(void);
/* ABI version (major * 100 + minor). */
int bdda_version(void);

/* FFT kernel selection for the whole process: 0 (default) uses the
 * two-step register kernels for the 5-smooth lengths m = R1*R2 with R1, R2
 * in {3,4,5,6,8,9,10,12,15,16} (32 lengths from 16 to 256), with the
 * dedicated x-pass kernel (forward x FFT, Green multiply, inverse x FFT in
 * one launch) where R1 is even on the compiled 8- and 4-column tiles, and
 * the runtime-radix LDS Stockham kernels for other lengths; 1 forces the
 * runtime-radix kernels everywhere; 2 is 0 without the dedicated x-pass
 * kernel (the generic register kernel runs the x pass).  Of the dedicated
 * x pass, 0 runs the split-spectrum kernel (two halves of the spectrum by
 * the parity of the first digit, half the LDS) where the radix pair is
 * balanced (R2 = R1 / 2: lengths 32, 50, 128) and the one-pass kernel
 * elsewhere; 3 runs the split kernel on every eligible length and 4 the
 * one-pass kernel everywhere.  1 to 4 exist to cross-check the kernels. */
int bdda_set_fft_path(int path);

/* Column groups of the occupied operator, for the whole process: 1 (default)
 * runs each FFT pass once over all column tiles; G (2..4) splits the tiles
 * into G groups, each on its own stream, started once the previous group's
 * y forward pass is done, so that one group's x pass overlaps the next
 * group's z and y passes.  The same arithmetic per column: results are
 * bit-identical to G = 1.  BDDA_E_ARG outside 1..4.  No reference
 * counterpart. */
int bdda_set_operator_groups(int groups);

/* Where the solvers' L x L recurrence algebra runs, for the whole process:
 * 0 (default) on the host after every Gram matrix (LAPACK-order LU with
 * partial pivoting, Cholesky, triangular inverse; 3-4 waits per iteration);
 * 1 on the device for L <= 64 (the same routines in single-workgroup
 * kernels; one wait per iteration, for the convergence test), on the host
 * for larger L.  On the MI355X box both take about the same time. */
int bdda_set_la_path(int path);

/* Plan over an nx x ny x nz cuboid lattice: chooses the convolution FFT
 * lengths m >= 2n - 1 per axis and owns every device buffer of the path.
 * Replaces the shape/allocation part of fft_init (mvp_fft/mvp_fft.py:97-119). */
int bdda_plan_create(int nx, int ny, int nz, bdda_plan** out);
int bdda_plan_destroy(bdda_plan* plan);
/* FFT lengths per axis (3 ints). */
int bdda_plan_fft_dims(const bdda_plan* plan, int* m_out);

/* Column tile Lb (orientations per line-FFT tile) the operator uses for L
 * columns: the FFT passes move ceil(L / Lb) * Lb columns.  Returns Lb, or
 * -BDDA_E_ARG.  For measurement (byte counts); no reference counterpart. */
int bdda_plan_column_tile(const bdda_plan* plan, int L);

/* Fix the plan's column tile (1 .. the widest tile its FFT lengths allow,
 * at most 16), or 0 to go back to the automatic choice above.  Returns
 * BDDA_E_ARG for a tile out of range.  No reference counterpart (the
 * reference's numpy FFTs have no tiles). */
int bdda_plan_set_column_tile(bdda_plan* plan, int lb);

/* The x pass kernel of the plan's operator (the pass that carries the
 * Green multiply, on the compiled column tiles): 0 follows the process-wide
 * rule of bdda_set_fft_path (the default), 1 the split-spectrum kernel on
 * every eligible length, 2 the one-pass kernel, 3 the split-spectrum kernel
 * where the radix pair is balanced (R2 = R1 / 2) and the one-pass one
 * elsewhere.  The products agree to rounding.  Returns BDDA_E_ARG for
 * other values.  No reference counterpart. */
int bdda_plan_set_x_kernel(bdda_plan* plan, int kernel);
/* The plan's x pass kernel setting (0-3 as above), or -BDDA_E_ARG. */
int bdda_plan_x_kernel(const bdda_plan* plan);

/* Autotune the column tile and the x pass kernel for L columns in
 * `precision`: times the occupied operator (one warm-up and three
 * applications per candidate, on `stream`) at the tiles 4, 8 and the
 * automatic choice, within the plan's limit, each with the split-spectrum
 * (1) and the one-pass (2) x pass, and fixes the fastest pair (the tile in
 * *lb_out, nullable; the kernel through bdda_plan_x_kernel).  Needs
 * bdda_green_init and bdda_set_target; allocates and frees one pair of
 * (3N, L) blocks.  The result is the same operator to rounding at any
 * setting.  For batched instances, whose lattices give every instance its
 * own FFT lengths; no reference counterpart. */
int bdda_plan_autotune(bdda_plan* plan, int L, int precision, int* lb_out, void* stream);

/* Green tensor A'(r) on the doubled grid and its 3-D FFT, kept on device in
 * complex128 and complex64 (6 unique components, half spectrum).
 * Replaces build_interaction_tensor + fft_init (mvp_fft/mvp_fft.py:13-119). */
int bdda_green_init(bdda_plan* plan, double lattice_lf, double k);
/* Share src's Green spectrum with dst (same cuboid): dst then computes with
 * it without building its own, and the spectrum lives until the last plan
 * holding it is destroyed.  A later bdda_green_init on either plan builds a
 * new spectrum for that plan only.  Lets one spectrum serve every
 * DiscreteDipoles of a (lattice, lattice_lf, k) -- the reference rebuilds
 * it per instance in set_interaction_matrix (bl_dda/scatterer.py:137-151),
 * although A_off does not depend on m_p; its sweep (run_dda.py:92-99)
 * varies m_p innermost. */
int bdda_green_share(bdda_plan* dst, const bdda_plan* src);

/* The full (m0, m1, m2, 3, 3) complex128 spectrum, i.e. what fft_init
 * returns when m = 2n on every axis (mvp_fft/mvp_fft.py:119). */
int bdda_green_download(const bdda_plan* plan, void* out_host);

/* The real-space (m0, m1, m2, 3, 3) complex128 tensor A'(r) before the FFT,
 * built on the device; equals build_interaction_tensor's array when m = 2n
 * on every axis.  Replaces build_interaction_tensor (mvp_fft/mvp_fft.py:13-94). */
int bdda_green_tensor(const bdda_plan* plan, double lattice_lf, double k, void* out_host);

/* Off-diagonal product on the full cuboid, Y = A_off P, for L columns.
 * P_dev, Y_dev: (3n, L) complex128.  precision: BDDA_PREC_FP64 or
 * BDDA_PREC_MIXED (complex64 transforms).
 * Replaces fft_mvp (mvp_fft/mvp_fft.py:122-173). */
int bdda_mvp_offdiag(bdda_plan* plan, const void* P_dev, void* Y_dev, int L, int precision, void* stream);

/* Occupied sites and the Jacobi-scaled diagonal.
 * address_host: (N) cuboid addresses (ascending, as np.where gives them);
 * pos_host: (N, 3) float64 positions (NULL: no positions; bdda_rhs and
 * bdda_observables then refuse); diag_host, jpre_host: (3N) complex128
 * diag(A) and 1/diag(A).
 * The plan also records the occupied extent (bounds of the occupied cells)
 * and the z lines holding an occupied cell: the occupied-block operator
 * (bdda_block_mvp, the solvers) skips every line and row outside them,
 * which hold only zeros.
 * Replaces _build_scatter_index and the jpre set-up of the solvers
 * (bl_krylov/bl_krylov.py:5-24,84,146). */
int bdda_set_target(bdda_plan* plan, const int64_t* address_host, int64_t N, const double* pos_host,
                    const void* diag_host, const void* jpre_host);

/* AS = jpre * (diag * S + A_off S) for L columns; S_dev, AS_dev: (3N, L)
 * complex128.  Replaces _block_mvp (bl_krylov/bl_krylov.py:27-55). */
int bdda_block_mvp(bdda_plan* plan, const void* S_dev, void* AS_dev, int L, void* stream);

/* Block-Krylov solve of A X = B with left Jacobi preconditioning.
 *   solver        BDDA_SOLVER_COCG_RQ | BDDA_SOLVER_BICGSTAB
 *   precision     BDDA_PREC_FP64: the reference algorithm in complex128;
 *                 BDDA_PREC_MIXED: the operator's Goodman convolution in
 *                 complex64 (complex128 scatter / gather boundaries and
 *                 Jacobi epilogue); Krylov blocks, updates, block QR and
 *                 L x L solves in float64.  Iterative refinement: each
 *                 cycle solves for a correction on the float64 true
 *                 residual.  Every refine_every iterations (0: never) a
 *                 complex64 cycle continues only while its residual is on
 *                 a log-linear course to tol by the next check and the
 *                 complex128 true residual is within 2x of the recursive
 *                 one.  A complex64 cycle that reaches its tolerance and at
 *                 least halves the true residual is followed, if needed, by
 *                 another complex64 cycle; otherwise the rest of the solve
 *                 runs complex128 cycles.
 *   B_dev         (3N, L) complex128 right-hand side
 *   X0_dev        (3N, L) complex128 initial guess or NULL (zero start); a
 *                 guess whose preconditioned residual exceeds ||jpre B|| is
 *                 discarded (zero start)
 *   X_dev         (3N, L) complex128 solution
 *   iter_fin      0-based index of the last iteration; err_fin the last
 *                 recurrence (FP64) or true (mixed) relative residual,
 *                 Frobenius over the block, / ||jpre B||
 *   true_res      recomputed ||jpre (B - A X)|| / ||jpre B|| (may be NULL)
 *   err_hist      per-iteration err (host, itermax doubles, may be NULL)
 *   flags_out     bit 0: the initial guess was used; bit 1 (mixed): the
 *                 solve fell back to complex128 cycles; bits 2-7 (mixed):
 *                 refinement cycles run (at most 63); bits 8 and up: the
 *                 number of err_hist entries written (may be NULL)
 * 1 <= L <= 4096 (the reference has no limit; larger blocks can be solved
 * in column groups).
 * Non-convergence is not an error: err_fin >= tol, as in the reference.
 * Replaces bl_cocg_rq_jacobi_mvp_fft and bl_bicgstab_jacobi_mvp_fft
 * (bl_krylov/bl_krylov.py:58-187). */
int bdda_solve(bdda_plan* plan, int solver, int precision, const void* B_dev, const void* X0_dev, void* X_dev,
               int L, double tol, int itermax, int refine_every, int* iter_fin, double* err_fin,
               double* true_res, double* err_hist, int* flags_out, void* stream);

/* True preconditioned relative residual ||jpre (B - A X)||_F / ||jpre B||_F
 * of a solution X_dev for B_dev, both (3N, L) complex128: one complex128
 * operator application and two norms (the stopping metric of
 * bl_krylov/bl_krylov.py:87-88,106,177, recomputed from X).  Host scalar
 * out; synchronises the stream. */
int bdda_true_residual(bdda_plan* plan, const void* B_dev, const void* X_dev, int L, double* res, void* stream);

/* Minimal-residual projection of an initial guess: X0_dev (3N, L)
 * complex128 is replaced by X0 C, where the L x L matrix C minimises
 * || jpre (B - A X0 C) ||_F over the span of the L guessed columns (one
 * complex128 operator application, two Gram matrices, an L x L solve).  The
 * result is never worse than the guess itself (C = I is admissible) nor than
 * zero (C = 0).  res_before / res_after (nullable): preconditioned relative
 * residual of the guess before and after.  No counterpart in the reference
 * (its solvers start from zero); meant for learned warm starts. */
int bdda_project_guess(bdda_plan* plan, const void* B_dev, void* X0_dev, int L, double* res_before,
                       double* res_after, void* stream);

/* Batched uneven instances: solves nb independent problems, instance i on
 * plans[i] (its own plan, with its own FFT lengths, column tile and Green
 * spectrum) with B[i] / X0[i] (nullable array or entries) / X[i] of shape
 * (3 N_i, L[i]) complex128 on the device.  nstreams host threads, each with
 * its own non-blocking HIP stream, take the instances largest first and run
 * bdda_solve on them concurrently; the per-instance results go to
 * iter_fin[i], err_fin[i], true_res[i], flags_out[i] and
 * err_hist[i * max(itermax, 1) ...] (each array nullable).  Returns after
 * every instance has finished; the first failing instance's error is
 * reported.  No counterpart in the reference (its solver takes one particle
 * at a time); a batch gives the same per-instance results as bdda_solve. */
int bdda_solve_batch(int nb, bdda_plan* const* plans, int solver, int precision, const void* const* B,
                     const void* const* X0, void* const* X, const int* L, double tol, int itermax, int refine_every,
                     int nstreams, int* iter_fin, double* err_fin, double* true_res, double* err_hist,
                     int* flags_out);

/* Per-pass kernel timing of the FFT operator with HIP events recorded on
 * the launch stream around every pass (z, y, x-convolution, inverse y,
 * inverse z; complex128 then complex64), for measurement only.
 * bdda_profile_read synchronises the device and returns, for the 10
 * (precision, pass) slots, the summed milliseconds and launch counts since
 * the last read, then resets them (at most 2048 launches are recorded). */
int bdda_profile_enable(bdda_plan* plan, int on);
int bdda_profile_read(bdda_plan* plan, double* ms_sum, int64_t* count);

/* Right-hand side B[3m+c][l] = exp(i k u_l . r_m) e0[l][c] into B_dev
 * (3N, L) complex128; u_host (L, 3) float64, e0_host (L, 3) complex128.
 * Replaces DiscreteDipoles.set_interaction_matrix :153-161 (bl_dda/scatterer.py). */
int bdda_rhs(bdda_plan* plan, const double* u_host, const void* e0_host, double k, int L, void* B_dev,
             void* stream);

/* Per-orientation observable sums from X_dev (3N, L) complex128.
 * geo_host (L, 24) float64: u_fw, th_fw, ph_fw, u_bk, th_bk, ph_bk, e0
 * (3 complex); epsfac_host (3) complex128: 4 pi / ((eps_r - 1) v) per axis.
 * out_host (L, 5) complex128:
 *   [0] sum P.th_fw conj(e_fw)       [1] sum P.ph_fw (-i) conj(e_fw)
 *   [2] sum P.th_bk conj(e_bk)       [3] sum P.ph_bk (-i) conj(e_bk)
 *   [4] (sum Im(P conj(E)), sum Im(P conj(E_inc)))
 * Replaces compute_C_abs / compute_C_ext / compute_PCAS_observable_S_fw /
 * compute_OCBS_observable_S_bk (bl_dda/scatterer.py:187-228). */
int bdda_observables(bdda_plan* plan, const void* X_dev, int L, const double* geo_host, const void* epsfac_host,
                     double k, void* out_host, void* stream);

/* bdda_observables without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Inputs and output are device pointers: geo_dev (L, 24)
 * float64 and epsfac_dev (3) complex128 as for bdda_observables, out_dev
 * (L, 5) complex128 receives the same sums (equal to bdda_observables').
 * Asynchronous on `stream`; calls no hipMalloc / hipFree / hipMemcpy.
 * The workgroup partials go to a fixed per-GPU scratch buffer of
 * 256 x 256 x 5 complex128 (5.2 MB), so 1 <= L <= 256 (else BDDA_E_ARG),
 * and calls that overlap on one GPU (other streams, other threads) race
 * on that buffer: serialise them. */
int bdda_observables_noalloc(bdda_plan* plan, const void* X_dev, int L, const double* geo_dev, const void* epsfac_dev,
                             double k, void* out_dev, void* stream);

/* Nearest surface-point distance of every lattice point, bit-exact with
 * the KD-tree query: sqrt(min_j ((dx*dx + dy*dy) + dz*dz)).
 * grid_host (n, 3), surf_host (P, 3) float64; dist_host (n) float64.
 * Replaces find_nearest_distance_from_the_GRE_surf
 * (shape_model/gaussian_ellipsoid.py:182-192). */
int bdda_nearest_distance(const double* grid_host, int64_t n, const double* surf_host, int64_t P,
                          double* dist_host);

/* Three-axis ray-cast inside test: near = dist < thresh; a voxel is inside
 * along an axis when a near voxel lies strictly before and strictly after
 * it; AND over the three axes.  dist_host (nx*ny*nz) float64, thresh =
 * distance_factor * lattice_lf; inside_host (nx*ny*nz) bytes (0/1).
 * Replaces extract_lattice_address_in_GRE_volume
 * (shape_model/gaussian_ellipsoid.py:195-237). */
int bdda_inside_mask(const double* dist_host, int nx, int ny, int nz, double thresh, uint8_t* inside_host);

/* ---- Learned warm start (no counterpart in the reference: the reference
 * has no learned component; these build the 3-D Fourier neural operator that
 * predicts an initial guess for bdda_solve's X0_dev).  float32 device
 * buffers. */

/* C = alpha * A B + beta * C for `batch` independent products with
 * arbitrary element strides: A[b*sAb + i*sAi + k*sAk] (M x K),
 * B[b*sBb + k*sBk + j*sBj] (K x N), C[b*sCb + i*sCi + j*sCj] (M x N).
 * Inputs are rounded to bfloat16 and multiplied on the MFMA units with
 * float32 accumulation. */
int bdda_gemm_bf16(int batch, int M, int N, int K, float alpha, const float* A, int64_t sAb, int64_t sAi,
                   int64_t sAk, const float* B, int64_t sBb, int64_t sBk, int64_t sBj, float beta, float* C,
                   int64_t sCb, int64_t sCi, int64_t sCj, void* stream);
/* The same product with an epilogue: bias (nullable, M floats) is added to
 * every row i of every product, then act = 1 applies GELU (act = 0: none).
 * Equal bit for bit to bdda_gemm_bf16 followed by bdda_bias_add and
 * bdda_gelu, without their passes over C.  BDDA_E_ARG for other act. */
int bdda_gemm_bf16_ex(int batch, int M, int N, int K, float alpha, const float* A, int64_t sAb, int64_t sAi,
                      int64_t sAk, const float* B, int64_t sBb, int64_t sBk, int64_t sBj, float beta, float* C,
                      int64_t sCb, int64_t sCi, int64_t sCj, const float* bias, int act, void* stream);
/* y = GELU(x) = 0.5 x (1 + erf(x / sqrt 2)); dx = dy * GELU'(x). */
int bdda_gelu(const float* x, float* y, int64_t n, void* stream);
int bdda_gelu_bwd(const float* x, const float* dy, float* dx, int64_t n, void* stream);
/* (B, C, S) tensors: y[b][c][s] += bias[c];  db[c] = sum_{b,s} dy[b][c][s]. */
int bdda_bias_add(float* y, const float* bias, int64_t B, int C, int64_t S, void* stream);
int bdda_bias_grad(const float* dy, float* db, int64_t B, int C, int64_t S, void* stream);
/* The network's y DFT as one wide product R = [F_re; F_im] Z (2K rows; Z's
 * nq slabs of 2K columns, K real parts then K imaginary parts) combined into
 * complex (nq, K, K) planes: xr[q][r][c] = R[r][q][c] - R[K+r][q][K+c],
 * xi[q][r][c] = R[r][q][K+c] + R[K+r][q][c], R row-major (2K, nq * 2K). */
int bdda_fno_ycombine(const float* R, float* xr, float* xi, int64_t nq, int K, void* stream);
/* Network input channels for nsamp samples on a G^3 cube (out: nsamp x 14
 * x G^3).  occ: per-instance occupancy cubes (bytes); sample_inst: instance
 * of each sample; desc: 20 floats per sample (chi re/im x3, u (3), e0 re/im
 * x3, k*lattice_lf, lattice origin / lattice_lf (3), unused). */
int bdda_fno_inputs(float* out, const uint8_t* occ, const int* sample_inst, const float* desc, int G, int64_t nsamp,
                    void* stream);
/* Between a (3N, L) complex128 block X and an (L, 6, G^3) float32 cube of
 * X / alpha (re, im per axis) at the given occupied cells: to_cube = 1
 * writes the cube (other cells untouched), 0 writes X = alpha * cube. */
int bdda_cube_exchange(float* cube, void* X, const int64_t* cells, const void* alpha, int64_t N, int L, int G,
                       int to_cube, void* stream);
/* Masked relative L2 loss per sample and (if grad != NULL) its gradient,
 * averaged over samples; pred/target (nsamp, C, S), occ (instances, S). */
int bdda_masked_l2(const float* pred, const float* target, const uint8_t* occ, const int* sample_inst, int C,
                   int64_t S, float* loss, float* grad, int nsamp, void* stream);
/* One Adam step (step counts from 1) on n parameters. */
int bdda_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps,
              int step, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* BDDA_H_ */
