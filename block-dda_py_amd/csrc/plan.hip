// Host side of libbdda: plans, the 5-launch Goodman FFT operator, the
// device-resident block-COCG-rq / block-BiCGSTAB solvers (float64 and
// mixed precision with float64 refinement) and the C ABI of include/bdda.h.
//
// Solver state lives on the device; per iteration only L x L Gram matrices
// (and a few scalars) travel to the host, where the float64 coefficient
// solves, Cholesky-QR factors and the convergence test run.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/bdda.h"
#include "host_la.h"
#include "internal.h"

namespace bdda {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

using la::zc;

#define BDDA_CHECK_LAUNCH()                                                                 \
  do {                                                                                      \
    hipError_t e_ = hipGetLastError();                                                      \
    if (e_ != hipSuccess) {                                                                 \
      set_error(std::string("kernel launch: ") + hipGetErrorString(e_));                    \
      return BDDA_E_HIP;                                                                    \
    }                                                                                       \
    if (!g_err.empty() && (g_err.rfind("line FFT", 0) == 0 || g_err.rfind("register FFT", 0) == 0)) \
      return BDDA_E_ARG;                                                                    \
  } while (0)

// Device buffer owned by a plan.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int ensure(size_t b) {
    if (b <= bytes) return 0;
    release();
    if (hipMalloc(&p, b) != hipSuccess) {
      (void)hipGetLastError();
      set_error("hipMalloc of " + std::to_string(b) + " bytes failed");
      return BDDA_E_OOM;
    }
    bytes = b;
    return 0;
  }
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Pinned host staging buffer.
struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
  int ensure(size_t b) {
    if (b <= bytes) return 0;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    if (hipHostMalloc(&p, b, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      set_error("hipHostMalloc failed");
      return BDDA_E_OOM;
    }
    bytes = b;
    return 0;
  }
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

}  // namespace bdda

struct bdda_plan {
  bdda::AxisPlan ax[3];
  long long ncell = 0;
  bool green_ready = false;
  bool target_ready = false;
  bool has_pos = false;
  long long N = 0;
  // occupied extent of the target: every occupied cell has ix < occ[0],
  // iy < occ[1], iz < occ[2] (the cuboid's dims until a target is set)
  int occ[3] = {0, 0, 0};
  // the z lines (ix * Ny + iy) that hold an occupied cell, ascending, and a
  // byte per line of the cuboid marking them
  bdda::DevBuf zlines, lmask;
  int nzl = 0;
  // twiddles
  bdda::DevBuf tw64[3], tw32[3];
  // Green half spectrum (complex128, complex64); shared between plans of
  // the same cuboid, lattice spacing and wavenumber (bdda_green_share)
  std::shared_ptr<bdda::DevBuf> g64, g32;
  double green_lf = 0.0, green_k = 0.0;
  // target
  bdda::DevBuf inv, pos, diag64, jpre64;
  // FFT intermediates (T1 per transform type: complex128, complex64)
  bdda::DevBuf T1, T1f, T2;
  // For the occupied operator, T1 holds zeros on every z line outside the
  // occupied-line list (the y forward pass reads them; nothing writes them).
  // t1_zero[type] records the layout that invariant holds for.
  struct T1Zero {
    void* ptr = nullptr;
    int Q = 0, nlo = 0;
    unsigned long long lines = 0;
    bool operator==(const T1Zero& o) const { return ptr == o.ptr && Q == o.Q && nlo == o.nlo && lines == o.lines; }
  } t1_zero[2];
  unsigned long long lines_version = 0;
  // solver workspace (blocks of 3N x L)
  bdda::DevBuf blk[12];
  bdda::DevBuf gram_part, gram_out, coef, dots_d, dots_n, obs_part, Bj64;
  bdda::HostBuf h_gram, h_coef, h_dots;
  // device L x L algebra: coefficient slots, and err / status / scalars /
  // pivots, with the pinned host copy of err and status read once an
  // iteration
  bdda::DevBuf dla, dla_misc;
  bdda::HostBuf h_sync;
  // column tile set by bdda_plan_set_column_tile / bdda_plan_autotune (0:
  // the rule of choose_lb)
  int lb_force = 0;
  // x pass kernel of this plan's operator (LinePass::xsplit): 0 the process-wide rule
  int x_kernel = 0;
  // optional per-pass kernel timing (HIP events recorded around each pass)
  bool profiling = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<int> ev_pass;  // pass id of each recorded pair (precision * 5 + pass)
  // overlapped operator: column groups 1.. run on their own streams, each
  // started once the previous group's y forward pass is done
  static constexpr int kMaxGroups = 4;
  hipStream_t grp_stream[kMaxGroups] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t grp_after_y[kMaxGroups] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t grp_done[kMaxGroups] = {nullptr, nullptr, nullptr, nullptr};
  ~bdda_plan() {
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    for (int g = 0; g < kMaxGroups; ++g) {
      if (grp_stream[g]) (void)hipStreamDestroy(grp_stream[g]);
      if (grp_after_y[g]) (void)hipEventDestroy(grp_after_y[g]);
      if (grp_done[g]) (void)hipEventDestroy(grp_done[g]);
    }
  }
};

namespace bdda {

static int check_plan(const bdda_plan* p) {
  if (!p) {
    set_error("null plan");
    return BDDA_E_ARG;
  }
  return 0;
}

// ------------------------------------------------------------------ operator
template <typename V> struct PrecTraits;
template <> struct PrecTraits<double2> {
  static const void* tw(const bdda_plan* p, int a) { return p->tw64[a].p; }
  static const void* green(const bdda_plan* p) { return p->g64->p; }
  static const void* diag(const bdda_plan* p) { return p->diag64.p; }
  static const void* jpre(const bdda_plan* p) { return p->jpre64.p; }
};
template <> struct PrecTraits<float2> {
  static const void* tw(const bdda_plan* p, int a) { return p->tw32[a].p; }
  static const void* green(const bdda_plan* p) { return p->g32->p; }
  // occupied blocks are always complex128 (see apply_operator)
  static const void* diag(const bdda_plan*) { return nullptr; }
  static const void* jpre(const bdda_plan*) { return nullptr; }
};

// Column tile Lb: every line-FFT tile (m x 3 Lb complex) must fit 4096
// values; among the admissible tiles pick the one that pads L least, then
// the 8-column tile (the register kernels have a compile-time version of it),
// then the widest (longer contiguous runs per grid point).
// Column groups of the occupied operator (bdda_set_operator_groups): 1 runs
// the five passes once over every column tile; G > 1 splits the tiles into
// G groups on staggered streams.
std::atomic<int> g_op_groups{1};

// Widest column tile every line-FFT kernel takes on this plan's lengths.
static int lb_cap(const bdda_plan* p) {
  const int mmax = std::max(p->ax[0].m, std::max(p->ax[1].m, p->ax[2].m));
  return std::max(1, std::min(4096 / (3 * mmax), 16));
}

static int choose_lb(const bdda_plan* p, int L) {
  if (p->lb_force > 0) return p->lb_force;
  const int cap = lb_cap(p);
  int best = 1;
  long long best_pad = (long long)L;
  for (int lb = cap; lb >= 1; --lb) {
    const long long pad = (long long)((L + lb - 1) / lb) * lb;
    if (pad < best_pad || (pad == best_pad && best != 8 && (lb == 8 || lb > best))) {
      best_pad = pad;
      best = lb;
    }
  }
  return best;
}

// Event pair around one launch when profiling is on.
static hipEvent_t prof_event(bdda_plan* p) {
  if (p->ev_used == p->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    p->ev_pool.push_back(e);
  }
  return p->ev_pool[p->ev_used++];
}

template <typename V>
static void launch_timed(bdda_plan* p, const LinePass& lp, int pass, hipStream_t s) {
  hipEvent_t a = nullptr, b = nullptr;
  if (p->profiling && p->ev_used + 2 <= 4096) {
    a = prof_event(p);
    b = prof_event(p);
  }
  if (a && b) (void)hipEventRecord(a, s);
  launch_line_pass<V>(lp, s);
  if (a && b) {
    (void)hipEventRecord(b, s);
    p->ev_pass.push_back((sizeof(V) == sizeof(float2) ? 5 : 0) + pass);
  }
}

template <typename V>
static int ensure_fft_ws(bdda_plan* p, int L) {
  const int Lb = choose_lb(p, L);
  const long long nlo = (L + Lb - 1) / Lb;
  const long long Q = 3LL * Lb;
  const long long nx = p->ax[0].n, ny = p->ax[1].n, mz = p->ax[2].m, my = p->ax[1].m;
  bdda::DevBuf& T1 = sizeof(V) == sizeof(double2) ? p->T1 : p->T1f;
  int rc = T1.ensure((size_t)(nlo * nx * ny * mz * Q) * sizeof(V));
  if (rc) return rc;
  return p->T2.ensure((size_t)(nlo * nx * my * mz * Q) * sizeof(V));
}

// Goodman convolution of L columns: z pass (scatter), y pass, fused x pass
// (forward, Green multiply, inverse), inverse y pass, inverse z pass (gather).
// occupied=true: S/AS are (3N, L) occupied blocks and the epilogue applies
// jpre * (diag * S + .); otherwise (3n, L) cuboid blocks and AS = A_off S.
// V is the transform's type; BV (occupied blocks only) the block's: a
// complex64 transform of complex128 blocks reads S and writes AS in
// complex128 (the epilogue in float64).
template <typename V, typename BV = V>
static int apply_operator(bdda_plan* p, const BV* S, BV* AS, int L, bool occupied, hipStream_t s) {
  constexpr bool wide = sizeof(BV) > sizeof(V);
  if ((occupied && sizeof(BV) != sizeof(double2)) || (wide && !occupied)) {
    set_error("occupied blocks are complex128 (cuboid blocks take the transform's type)");
    return BDDA_E_ARG;
  }
  const int Lb = choose_lb(p, L);
  const int Q = 3 * Lb;
  const int nlo = (L + Lb - 1) / Lb;
  const int nx = p->ax[0].n, ny = p->ax[1].n, nz = p->ax[2].n;
  const int mx = p->ax[0].m, my = p->ax[1].m, mz = p->ax[2].m;
  // Occupied blocks: cells beyond the target's occupied extent are empty, so
  // the lines and rows there hold only zeros and every pass skips them (a
  // cuboid padded beyond the lattice costs nothing); x runs to ox, y to oy
  // and z to oz.  T1 lines and T2 rows outside are neither written nor read.
  const int ox = occupied ? p->occ[0] : nx, oy = occupied ? p->occ[1] : ny, oz = occupied ? p->occ[2] : nz;
  if (ox < 1 || oy < 1 || oz < 1 || ox > nx || oy > ny || oz > nz) {
    set_error("apply_operator: occupied extent outside the cuboid");
    return BDDA_E_STATE;
  }
  int rc = ensure_fft_ws<V>(p, L);
  if (rc) return rc;
  g_err.clear();
  const long long t1_lo = (long long)nx * ny * mz * Q;
  const long long t2_lo = (long long)nx * my * mz * Q;
  bdda::DevBuf& T1 = sizeof(V) == sizeof(double2) ? p->T1 : p->T1f;
  {
    // zeros on the z lines the occupied operator skips (once per layout);
    // the cuboid passes write every line
    auto& z = p->t1_zero[sizeof(V) == sizeof(double2) ? 0 : 1];
    if (occupied) {
      bdda_plan::T1Zero key;
      key.ptr = T1.p;
      key.Q = Q;
      key.nlo = nlo;
      key.lines = p->lines_version;
      if (!(z == key)) {
        BDDA_HIP(hipMemsetAsync(T1.p, 0, (size_t)(nlo * t1_lo) * sizeof(V), s));
        z = key;
      }
    } else {
      z = bdda_plan::T1Zero();
    }
  }
  LinePass base;
  std::memset(&base, 0, sizeof(base));
  base.Q = Q;
  base.L = L;
  base.Lb = Lb;
  base.nlo = nlo;
  base.nz = nz;
  base.nyz_y = ny;
  base.inv = p->inv.as<int>();
  base.S = S;
  base.AS = AS;
  base.diag = PrecTraits<BV>::diag(p);
  base.jpre = PrecTraits<BV>::jpre(p);
  base.mx = mx;
  base.my = my;
  base.mz = mz;
  base.hx = p->ax[0].h;
  base.hy = p->ax[1].h;
  base.hz = p->ax[2].h;
  base.green = PrecTraits<V>::green(p);
  auto axis = [&](LinePass& lp, int a) {
    lp.m = p->ax[a].m;
    lp.nrad = p->ax[a].nrad;
    for (int i = 0; i < kMaxRadix; ++i) lp.rad[i] = p->ax[a].rad[i];
    lp.tw = PrecTraits<V>::tw(p, a);
  };
  LinePass lps[5];
  // P1: z forward with scatter (points (ix, iy))
  {
    LinePass lp = base;
    axis(lp, 2);
    lp.mode = occupied ? (wide ? PASS_SCATTER_Z64 : PASS_SCATTER_Z) : PASS_CUBOID_Z;
    lp.nin = oz;
    lp.nout = mz;
    lp.npts = nx * ny;  // line pt = ix * Ny + iy
    lp.ptb = ny;
    if (occupied) {  // only the lines that hold an occupied cell
      lp.zlines = p->zlines.as<int>();
      lp.npts = p->nzl;
    }
    lp.da = (long long)ny * mz * Q;
    lp.db = (long long)mz * Q;
    lp.de = Q;
    lp.lo_dst = t1_lo;
    lp.dst = T1.p;
    lp.order_pe = 0;
    lps[0] = lp;
  }
  // P2: y forward (points (ix, kz)), T1 -> T2
  {
    LinePass lp = base;
    axis(lp, 1);
    lp.mode = PASS_PLAIN;
    lp.nin = oy;
    lp.nout = my;
    lp.npts = ox * mz;
    lp.ptb = mz;
    lp.sa = (long long)ny * mz * Q;
    lp.sb = Q;
    lp.se = (long long)mz * Q;
    lp.lo_src = t1_lo;
    lp.da = (long long)my * mz * Q;
    lp.db = Q;
    lp.de = (long long)mz * Q;
    lp.lo_dst = t2_lo;
    lp.src = T1.p;
    lp.dst = p->T2.p;
    lp.order_pe = 1;
    lps[1] = lp;
  }
  // P3: x forward * Green * x inverse (points (ky, kz)), in place on T2
  {
    LinePass lp = base;
    axis(lp, 0);
    lp.mode = PASS_CONV_X;
    lp.xsplit = p->x_kernel;
    lp.nin = ox;
    lp.nout = ox;
    lp.npts = my * mz;
    lp.ptb = lp.npts;
    lp.sb = Q;
    lp.se = (long long)my * mz * Q;
    lp.db = Q;
    lp.de = (long long)my * mz * Q;
    lp.lo_src = t2_lo;
    lp.lo_dst = t2_lo;
    lp.src = p->T2.p;
    lp.dst = p->T2.p;
    lp.order_pe = 1;
    lps[2] = lp;
  }
  // P4: y inverse (points (ix, kz)), T2 -> T1
  {
    LinePass lp = base;
    axis(lp, 1);
    lp.mode = PASS_PLAIN;
    lp.inverse = 1;
    lp.nin = my;
    lp.nout = oy;
    lp.npts = ox * mz;
    if (occupied) lp.lmask = p->lmask.as<unsigned char>();  // T1 keeps zeros off the occupied lines
    lp.ptb = mz;
    lp.sa = (long long)my * mz * Q;
    lp.sb = Q;
    lp.se = (long long)mz * Q;
    lp.lo_src = t2_lo;
    lp.da = (long long)ny * mz * Q;
    lp.db = Q;
    lp.de = (long long)mz * Q;
    lp.lo_dst = t1_lo;
    lp.src = p->T2.p;
    lp.dst = T1.p;
    lp.order_pe = 1;
    lps[3] = lp;
  }
  // P5: z inverse with crop and gather/epilogue (points (ix, iy))
  {
    LinePass lp = base;
    axis(lp, 2);
    lp.mode = occupied ? (wide ? PASS_GATHER_Z64 : PASS_GATHER_Z) : PASS_CROP_Z;
    lp.inverse = 1;
    lp.nin = mz;
    lp.nout = oz;
    lp.npts = nx * ny;
    lp.ptb = ny;
    if (occupied) {
      lp.zlines = p->zlines.as<int>();
      lp.npts = p->nzl;
    }
    lp.sa = (long long)ny * mz * Q;
    lp.sb = (long long)mz * Q;
    lp.se = Q;
    lp.lo_src = t1_lo;
    lp.src = T1.p;
    lp.scale = 1.0 / ((double)mx * my * mz);
    lp.order_pe = 0;
    lps[4] = lp;
  }
  // The passes of column tiles [t0, t1) on stream st (the T1 / T2 blocks of
  // those tiles, and their columns of S and AS); after_y (nullable) is
  // recorded once their y forward pass is queued.
  auto run_tiles = [&](int t0, int t1, hipStream_t st, hipEvent_t after_y) -> int {
    for (int k = 0; k < 5; ++k) {
      LinePass lp = lps[k];
      lp.nlo = t1 - t0;
      lp.col0 = t0 * Lb;
      if (lp.src) lp.src = reinterpret_cast<const V*>(lp.src) + (long long)t0 * lp.lo_src;
      if (lp.dst) lp.dst = reinterpret_cast<V*>(lp.dst) + (long long)t0 * lp.lo_dst;
      launch_timed<V>(p, lp, k, st);
      BDDA_CHECK_LAUNCH();
      if (k == 1 && after_y) BDDA_HIP(hipEventRecord(after_y, st));
    }
    return 0;
  };
  // Column groups: the x pass of one group (in-core bound) overlaps the
  // z and y passes of the next (HBM bound).  Per-pass profiling runs the
  // plain sequence.
  const int G = occupied && !p->profiling ? std::min(std::min(g_op_groups.load(), nlo), bdda_plan::kMaxGroups) : 1;
  if (G <= 1) return run_tiles(0, nlo, s, nullptr);
  for (int g = 0; g < G; ++g) {
    if (!p->grp_after_y[g]) BDDA_HIP(hipEventCreateWithFlags(&p->grp_after_y[g], hipEventDisableTiming));
    if (!p->grp_done[g]) BDDA_HIP(hipEventCreateWithFlags(&p->grp_done[g], hipEventDisableTiming));
    if (g > 0 && !p->grp_stream[g]) BDDA_HIP(hipStreamCreateWithFlags(&p->grp_stream[g], hipStreamNonBlocking));
  }
  for (int g = 0; g < G; ++g) {
    const int t0 = (int)((long long)nlo * g / G), t1 = (int)((long long)nlo * (g + 1) / G);
    hipStream_t st = g == 0 ? s : p->grp_stream[g];
    if (g > 0) BDDA_HIP(hipStreamWaitEvent(st, p->grp_after_y[g - 1], 0));
    if ((rc = run_tiles(t0, t1, st, p->grp_after_y[g]))) return rc;
    if (g > 0) BDDA_HIP(hipEventRecord(p->grp_done[g], st));
  }
  for (int g = 1; g < G; ++g) BDDA_HIP(hipStreamWaitEvent(s, p->grp_done[g], 0));
  return 0;
}

// ------------------------------------------------------------ solver helpers
struct Ctx {
  bdda_plan* p;
  int L;
  long long rows;
  hipStream_t s;
  std::vector<double>* hist;
  // mixed mode: every check_every iterations the true residual is computed
  // with the complex128 operator; once it exceeds twice the recursive one
  // (or is not finite) the solve stops with conv32 cleared, and the caller
  // restarts from the true residual with the complex128 convolution
  int check_every = 0;
  // mixed mode: the operator's convolution runs in complex64 (complex128
  // blocks at its scatter / gather boundaries)
  bool conv32 = false;
};

// Gram matrices op(A)^H B (conj) or A^T B into device / pinned host slot
// `slot` (0..2) without waiting: several Grams (and a norm) then share one
// stream synchronisation, gram_unpack reads a slot after host_wait.
constexpr int kGramSlots = 3;
static int gram_launch(Ctx& c, const void* A, const void* B, bool conj, bool f32, int slot) {
  bdda_plan* p = c.p;
  const int L = c.L;
  const int ntile = (L + 63) / 64;
  const int np = gram_parts(L, c.rows);
  const size_t one = (size_t)L * L * sizeof(double2);
  int rc = p->gram_part.ensure((size_t)np * ntile * ntile * 4096 * sizeof(double2));
  if (!rc) rc = p->gram_out.ensure(kGramSlots * one);
  if (!rc) rc = p->h_gram.ensure(kGramSlots * one);
  if (rc) return rc;
  double2* dout = p->gram_out.as<double2>() + (size_t)slot * L * L;
  if (f32)
    launch_gram<float2>((const float2*)A, (const float2*)B, c.rows, L, conj, p->gram_part.as<double2>(), np, dout,
                        c.s);
  else
    launch_gram<double2>((const double2*)A, (const double2*)B, c.rows, L, conj, p->gram_part.as<double2>(), np,
                         dout, c.s);
  BDDA_HIP(hipGetLastError());
  BDDA_HIP(hipMemcpyAsync(p->h_gram.as<char>() + slot * one, dout, one, hipMemcpyDeviceToHost, c.s));
  return 0;
}

static int host_wait(Ctx& c) {
  BDDA_HIP(hipStreamSynchronize(c.s));
  return 0;
}

static void gram_unpack(Ctx& c, int slot, std::vector<zc>& out) {
  const size_t one = (size_t)c.L * c.L * sizeof(double2);
  out.resize((size_t)c.L * c.L);
  std::memcpy(out.data(), c.p->h_gram.as<char>() + slot * one, one);
}

static int gram_to_host(Ctx& c, const void* A, const void* B, bool conj, bool f32, std::vector<zc>& out) {
  int rc = gram_launch(c, A, B, conj, f32, 0);
  if (!rc) rc = host_wait(c);
  if (rc) return rc;
  gram_unpack(c, 0, out);
  return 0;
}

// Upload an L x L coefficient matrix into coefficient slot `slot` (pinned
// staging; every reuse of a slot is separated by a stream synchronisation in
// the solvers, so the host never overwrites a slot still being copied).
static int coef_upload(Ctx& c, const std::vector<zc>& C, int slot, const double2** dev) {
  bdda_plan* p = c.p;
  const size_t one = (size_t)c.L * c.L * sizeof(double2);
  int rc = p->coef.ensure(8 * one);
  if (!rc) rc = p->h_coef.ensure(8 * one);
  if (rc) return rc;
  char* h = p->h_coef.as<char>() + slot * one;
  std::memcpy(h, C.data(), one);
  char* d = p->coef.as<char>() + slot * one;
  BDDA_HIP(hipMemcpyAsync(d, h, one, hipMemcpyHostToDevice, c.s));
  *dev = reinterpret_cast<const double2*>(d);
  return 0;
}

template <typename V>
static int dots_to_host(Ctx& c, const V* A, const V* B, zc* dot, double* nrm) {
  bdda_plan* p = c.p;
  const int np = dot_parts();
  int rc = p->dots_d.ensure(np * sizeof(double2));
  if (!rc) rc = p->dots_n.ensure(np * sizeof(double));
  if (!rc) rc = p->h_dots.ensure(np * (sizeof(double2) + sizeof(double)));
  if (rc) return rc;
  launch_dot<V>(A, B, c.rows * c.L, dot ? p->dots_d.as<double2>() : nullptr, nrm ? p->dots_n.as<double>() : nullptr,
                np, c.s);
  BDDA_HIP(hipGetLastError());
  double2* hd = p->h_dots.as<double2>();
  double* hn = reinterpret_cast<double*>(hd + np);
  if (dot) BDDA_HIP(hipMemcpyAsync(hd, p->dots_d.p, np * sizeof(double2), hipMemcpyDeviceToHost, c.s));
  if (nrm) BDDA_HIP(hipMemcpyAsync(hn, p->dots_n.p, np * sizeof(double), hipMemcpyDeviceToHost, c.s));
  BDDA_HIP(hipStreamSynchronize(c.s));
  if (dot) {
    zc s = 0.0;
    for (int i = 0; i < np; ++i) s += zc(hd[i].x, hd[i].y);
    *dot = s;
  }
  if (nrm) {
    double s = 0.0;
    for (int i = 0; i < np; ++i) s += hn[i];
    *nrm = s;
  }
  return 0;
}

// Y = X + a Z with |Y|_F^2 partial sums copied to the host (no wait);
// axpy_norm_sum adds them up after host_wait.
template <typename V>
static int axpy_norm_launch(Ctx& c, V* Y, const V* X, zc a, const V* Z) {
  bdda_plan* p = c.p;
  const int np = dot_parts();
  int rc = p->dots_n.ensure(np * sizeof(double));
  if (!rc) rc = p->h_dots.ensure(np * (sizeof(double2) + sizeof(double)));
  if (rc) return rc;
  launch_axpy_norm<V>(Y, X, make_double2(a.real(), a.imag()), Z, c.rows * c.L, p->dots_n.as<double>(), np, c.s);
  BDDA_HIP(hipGetLastError());
  BDDA_HIP(hipMemcpyAsync(p->h_dots.p, p->dots_n.p, np * sizeof(double), hipMemcpyDeviceToHost, c.s));
  return 0;
}

static double axpy_norm_sum(Ctx& c) {
  const double* hn = c.p->h_dots.as<double>();
  double s = 0.0;
  for (int i = 0; i < dot_parts(); ++i) s += hn[i];
  return s;
}

template <typename V>
static int axpy_norm_host(Ctx& c, V* Y, const V* X, zc a, const V* Z, double* nrm2) {
  int rc = axpy_norm_launch<V>(c, Y, X, a, Z);
  if (!rc) rc = host_wait(c);
  if (rc) return rc;
  *nrm2 = axpy_norm_sum(c);
  return 0;
}

template <typename V>
static int update(Ctx& c, V* Y, const V* X, zc a, const V* W1, const V* W2, zc w2c, const std::vector<zc>& C,
                  int slot, const V* E, zc e) {
  const double2* dC = nullptr;
  int rc = coef_upload(c, C, slot, &dC);
  if (rc) return rc;
  launch_update<V>(Y, X, make_double2(a.real(), a.imag()), W1, W2, make_double2(w2c.real(), w2c.imag()), dC, E,
                   make_double2(e.real(), e.imag()), c.rows, c.L, c.s);
  BDDA_HIP(hipGetLastError());
  return 0;
}

// The update Y = a X + W1 C on the matrix cores with the Gram matrix of Y
// (Y^H Y when conj, else Y^T Y) into pinned host slot `gslot` without
// waiting (complex128 blocks, L <= kUpdateGramMaxL): the Gram needs no second
// pass over Y.
static int update_gram_launch(Ctx& c, double2* Y, const double2* X, zc a, const double2* W1,
                              const std::vector<zc>& C, int cslot, bool conj, int gslot) {
  bdda_plan* p = c.p;
  const int L = c.L;
  const size_t one = (size_t)L * L * sizeof(double2);
  const double2* dC = nullptr;
  int rc = coef_upload(c, C, cslot, &dC);
  if (!rc) rc = p->gram_part.ensure((size_t)update_gram_parts(c.rows) * 32 * 32 * sizeof(double2));
  if (!rc) rc = p->gram_out.ensure(kGramSlots * one);
  if (!rc) rc = p->h_gram.ensure(kGramSlots * one);
  if (rc) return rc;
  double2* dout = p->gram_out.as<double2>() + (size_t)gslot * L * L;
  launch_update_gram(Y, X, make_double2(a.real(), a.imag()), W1, nullptr, make_double2(0.0, 0.0), dC, nullptr,
                     make_double2(0.0, 0.0), c.rows, L, conj ? 1 : 2, p->gram_part.as<double2>(), dout, c.s);
  BDDA_HIP(hipGetLastError());
  BDDA_HIP(hipMemcpyAsync(p->h_gram.as<char>() + gslot * one, dout, one, hipMemcpyDeviceToHost, c.s));
  return 0;
}

template <typename V> static constexpr bool is_f32() { return sizeof(V) == sizeof(float2); }
template <typename V> static bool fused_gram(const Ctx& c) {
  return std::is_same<V, double2>::value && c.L <= kUpdateGramMaxL;
}

// Mixed mode, at iteration k of a complex64-convolution cycle: keep going
// only while the recursive residual err is on a log-linear course to tol by
// the next check (err <= tol^(k / (k + every))); a solve that is not is
// hard, and complex64 rounding in its products delays convergence for good
// (lost conjugacy), so the caller restarts it in complex128.
static bool conv32_on_track(double err, double tol, int k, int every) {
  return err <= std::pow(tol, (double)k / (double)(k + every));
}

template <typename V>
static int op(Ctx& c, const V* S, V* AS) {
  if constexpr (std::is_same<V, double2>::value) {
    if (c.conv32) return apply_operator<float2, double2>(c.p, S, AS, c.L, true, c.s);
  }
  return apply_operator<V>(c.p, S, AS, c.L, true, c.s);
}

static int singular(const char* what) {
  set_error(std::string("singular matrix in ") + what);
  return BDDA_E_SINGULAR;
}

// Block-BiCGSTAB on the preconditioned system (bl_krylov.py:121-187).
// Bj: preconditioned right-hand side; X: in/out (x_zero -> starts from 0).
// bnorm: ||jpre B||_F used for err.  ws: 7 work blocks.
template <typename V>
static int solve_bicgstab(Ctx& c, const V* Bj, V* X, bool x_zero, double bnorm, double tol, int itermax, V** ws,
                          int* it_out, double* err_out) {
  const long long n = c.rows * c.L;
  const size_t bytes = n * sizeof(V);
  V *R = ws[0], *P = ws[1], *R0 = ws[2], *Vb = ws[3], *T = ws[4], *Z = ws[5], *Pn = ws[6];
  int rc;
  if (x_zero) {
    BDDA_HIP(hipMemsetAsync(X, 0, bytes, c.s));
    BDDA_HIP(hipMemcpyAsync(R, Bj, bytes, hipMemcpyDeviceToDevice, c.s));
  } else {
    if ((rc = op<V>(c, X, T))) return rc;
    double dummy;
    if ((rc = axpy_norm_host<V>(c, R, Bj, zc(-1.0), T, &dummy))) return rc;
  }
  BDDA_HIP(hipMemcpyAsync(P, R, bytes, hipMemcpyDeviceToDevice, c.s));
  BDDA_HIP(hipMemcpyAsync(R0, R, bytes, hipMemcpyDeviceToDevice, c.s));
  const int L = c.L;
  std::vector<zc> RV, RR, RZ, alpha, beta, negA;
  std::vector<int> piv;
  int it = 0;
  double err = INFINITY;
  for (int k = 0; k < itermax; ++k) {
    if constexpr (std::is_same<V, double2>::value) {
      if (c.conv32 && c.check_every > 0 && k > 0 && k % c.check_every == 0) {
        // stop (the caller restarts in complex128) when the solve is off
        // course, or when the true residual, from the complex128 operator,
        // has drifted from the recursive one
        bool stop = !conv32_on_track(err, tol, k, c.check_every);
        if (!stop) {
          if ((rc = apply_operator<double2>(c.p, X, T, L, true, c.s))) return rc;
          double rt2;
          if ((rc = axpy_norm_host<V>(c, Z, Bj, zc(-1.0), T, &rt2))) return rc;
          stop = !(std::sqrt(rt2) / bnorm <= 2.0 * err);
        }
        if (stop) {
          c.conv32 = false;
          break;
        }
      }
    }
    if ((rc = op<V>(c, P, Vb))) return rc;
    if ((rc = gram_launch(c, R0, Vb, true, is_f32<V>(), 0)) || (rc = gram_launch(c, R0, R, true, is_f32<V>(), 1)) ||
        (rc = host_wait(c)))
      return rc;
    gram_unpack(c, 0, RV);
    gram_unpack(c, 1, RR);
    std::vector<zc> LU = RV;
    if (!la::lu_factor(LU, piv, L)) return singular("block-BiCGSTAB (R0^H V)");
    alpha = RR;
    la::lu_solve(LU, piv, L, alpha, L);
    negA.resize(alpha.size());
    for (size_t i = 0; i < alpha.size(); ++i) negA[i] = -alpha[i];
    // T = R - V alpha
    if ((rc = update<V>(c, T, R, 1.0, Vb, nullptr, 0.0, negA, 0, nullptr, 0.0))) return rc;
    if ((rc = op<V>(c, T, Z))) return rc;
    // xi = tr(Z^H T) / tr(Z^H Z) from one pass: sum conj(T) Z and |Z|^2
    zc tz;
    double zn2;
    if ((rc = dots_to_host<V>(c, T, Z, &tz, &zn2))) return rc;
    const zc xi = std::conj(tz) / zn2;
    // X = X + P alpha + xi T
    if ((rc = update<V>(c, X, X, 1.0, P, nullptr, 0.0, alpha, 1, T, xi))) return rc;
    // R = T - xi Z with its norm, and R0^H Z for the next beta, one wait
    if ((rc = axpy_norm_launch<V>(c, R, T, -xi, Z)) || (rc = gram_launch(c, R0, Z, true, is_f32<V>(), 2)) ||
        (rc = host_wait(c)))
      return rc;
    const double rn2 = axpy_norm_sum(c);
    err = std::sqrt(rn2) / bnorm;
    if (c.hist) c.hist->push_back(err);
    it = k;
    if (err < tol) break;
    gram_unpack(c, 2, RZ);
    beta = RZ;
    for (auto& v : beta) v = -v;
    la::lu_solve(LU, piv, L, beta, L);
    // Pn = R + (P - xi V) beta
    if ((rc = update<V>(c, Pn, R, 1.0, P, Vb, -xi, beta, 2, nullptr, 0.0))) return rc;
    std::swap(P, Pn);
  }
  *it_out = it;
  *err_out = err;
  return 0;
}

// The block QR's last resort: Householder QR of W on the host (LAPACK's
// reflectors, as numpy.linalg.qr in the reference), Q into Qn, R into tau.
// Taken when even shifted CholeskyQR3 breaks down, i.e. W is numerically
// rank-deficient, as a block of more orientations than the particle's field
// has independent modes becomes; the reference's QR proceeds there too.
template <typename V>
static int householder_host(Ctx& c, const V* W, V* Qn, std::vector<zc>& tau) {
  const long long rows = c.rows;
  const int L = c.L;
  const size_t n = (size_t)rows * L;
  std::vector<V> h(n);
  BDDA_HIP(hipMemcpyAsync(h.data(), W, n * sizeof(V), hipMemcpyDeviceToHost, c.s));
  BDDA_HIP(hipStreamSynchronize(c.s));
  std::vector<double> re(n), im(n);
  for (long long r = 0; r < rows; ++r)
    for (int l = 0; l < L; ++l) {
      re[(size_t)l * rows + r] = h[(size_t)r * L + l].x;
      im[(size_t)l * rows + r] = h[(size_t)r * L + l].y;
    }
  la::householder_qr(re, im, rows, L, tau);
  using T = decltype(h[0].x);
  for (long long r = 0; r < rows; ++r)
    for (int l = 0; l < L; ++l) {
      h[(size_t)r * L + l].x = (T)re[(size_t)l * rows + r];
      h[(size_t)r * L + l].y = (T)im[(size_t)l * rows + r];
    }
  BDDA_HIP(hipMemcpyAsync(Qn, h.data(), n * sizeof(V), hipMemcpyHostToDevice, c.s));
  BDDA_HIP(hipStreamSynchronize(c.s));
  return 0;
}

// Reduced QR of W by Cholesky-QR2 in float64: Qn (device) and tau (host).
template <typename V>
static int cholqr2(Ctx& c, const V* W, V* Q1, V* Qn, std::vector<zc>& tau) {
  const int L = c.L;
  std::vector<zc> G, Ri, R1, R2, R3, t;
  int rc = 0;
  // one Cholesky QR pass: out = in R^-1 with R^H R = in^H in + shift I
  // (*ok = false when the Cholesky factorisation breaks down)
  auto pass = [&](const V* in, V* out, std::vector<zc>& R, bool shifted, int slot, bool* ok) -> int {
    int e = gram_to_host(c, in, in, true, is_f32<V>(), G);
    if (e) return e;
    if (shifted) {
      // shifted CholeskyQR: s = 11 (m n + n (n + 1)) u tr(G) makes G + s I
      // positive definite for any block with cond < 1/u
      double tr = 0.0;
      for (int i = 0; i < L; ++i) tr += G[(size_t)i * L + i].real();
      const double shift = 11.0 * ((double)c.rows * L + (double)L * (L + 1)) * std::ldexp(1.0, -53) * tr;
      for (int i = 0; i < L; ++i) G[(size_t)i * L + i] += shift;
    }
    *ok = la::cholesky_upper(G, R, L);
    if (!*ok) return 0;
    la::upper_inverse(R, Ri, L);
    return update<V>(c, out, nullptr, 0.0, in, nullptr, 0.0, Ri, slot, nullptr, 0.0);
  };
  bool ok1 = false, ok2 = false, ok3 = false;
  if ((rc = pass(W, Q1, R1, false, 3, &ok1))) return rc;
  if (ok1 && (rc = pass(Q1, Qn, R2, false, 4, &ok2))) return rc;
  if (ok1 && ok2) {
    la::matmul(R2, false, R1, tau, L);
    return 0;
  }
  // Nearly rank-deficient block (the Gram matrix squares its condition
  // number): shifted CholeskyQR3 from W -- a shifted pass, then two plain
  // passes that restore orthogonality; W is left intact for the last resort.
  if ((rc = pass(W, Q1, R1, true, 3, &ok1))) return rc;
  if (ok1 && (rc = pass(Q1, Qn, R2, false, 4, &ok2))) return rc;
  if (ok1 && ok2 && (rc = pass(Qn, Q1, R3, false, 4, &ok3))) return rc;
  if (!(ok1 && ok2 && ok3)) return householder_host<V>(c, W, Qn, tau);
  BDDA_HIP(hipMemcpyAsync(Qn, Q1, (size_t)c.rows * L * sizeof(V), hipMemcpyDeviceToDevice, c.s));
  la::matmul(R2, false, R1, t, L);
  la::matmul(R3, false, t, tau, L);
  return 0;
}

// Cholesky-QR2 of W given its Gram matrix G1 = W^H W (from the update that
// wrote W), for complex128 blocks with L <= kUpdateGramMaxL: the two
// passes' updates Q1 = W R1^-1 and Qn = Q1 R2^-1 produce the Gram matrices
// of their outputs, Q1^H Q1 for the second pass and Qn^T Qn (the next beta's)
// into host slot 1, launched but not waited for (*qnqn_pending).  When a
// Cholesky factorisation fails, cholqr2 (shifted CholeskyQR3, then
// Householder) takes over from W and Qn^T Qn is left to the caller.
static int cholqr2_fused(Ctx& c, const double2* W, const std::vector<zc>& G1, double2* Q1, double2* Qn,
                         std::vector<zc>& tau, bool* qnqn_pending) {
  const int L = c.L;
  std::vector<zc> R1, R2, Ri, G2;
  *qnqn_pending = false;
  int rc;
  if (la::cholesky_upper(G1, R1, L)) {
    la::upper_inverse(R1, Ri, L);
    if ((rc = update_gram_launch(c, Q1, nullptr, 0.0, W, Ri, 3, true, 0)) || (rc = host_wait(c))) return rc;
    gram_unpack(c, 0, G2);
    if (la::cholesky_upper(G2, R2, L)) {
      la::upper_inverse(R2, Ri, L);
      if ((rc = update_gram_launch(c, Qn, nullptr, 0.0, Q1, Ri, 4, false, 1))) return rc;
      *qnqn_pending = true;
      la::matmul(R2, false, R1, tau, L);
      return 0;
    }
  }
  return cholqr2<double2>(c, W, Q1, Qn, tau);
}

// Block-COCG-rq on the preconditioned system (bl_krylov.py:58-118).
// Complex128 blocks with L <= kUpdateGramMaxL take the Gram matrices of the
// block QR and of the next beta from the updates that write those blocks.
template <typename V>
static int solve_cocg(Ctx& c, const V* Bj, V* X, bool x_zero, double bnorm, double tol, int itermax, V** ws,
                      int* it_out, double* err_out) {
  const long long n = c.rows * c.L;
  const size_t bytes = n * sizeof(V);
  V *Q = ws[0], *S = ws[1], *AS = ws[2], *W = ws[3], *Q1 = ws[4], *Sn = ws[5], *R0 = ws[6];
  int rc;
  const int L = c.L;
  if (x_zero) {
    BDDA_HIP(hipMemsetAsync(X, 0, bytes, c.s));
    BDDA_HIP(hipMemcpyAsync(R0, Bj, bytes, hipMemcpyDeviceToDevice, c.s));
  } else {
    if ((rc = op<V>(c, X, AS))) return rc;
    double dummy;
    if ((rc = axpy_norm_host<V>(c, R0, Bj, zc(-1.0), AS, &dummy))) return rc;
  }
  std::vector<zc> xi, QtQ, StAS, alpha, ax, tau, QntQn, rhs, beta, negA;
  std::vector<int> piv;
  if ((rc = cholqr2<V>(c, R0, Q1, Q, xi))) return rc;
  BDDA_HIP(hipMemcpyAsync(S, Q, bytes, hipMemcpyDeviceToDevice, c.s));
  if ((rc = gram_to_host(c, Q, Q, false, is_f32<V>(), QtQ))) return rc;
  int it = 0;
  double err = INFINITY;
  for (int k = 0; k < itermax; ++k) {
    if constexpr (std::is_same<V, double2>::value) {
      if (c.conv32 && c.check_every > 0 && k > 0 && k % c.check_every == 0) {
        // stop (the caller restarts in complex128) when the solve is off
        // course, or when the true residual, from the complex128 operator,
        // has drifted from the recursive one
        bool stop = !conv32_on_track(err, tol, k, c.check_every);
        if (!stop) {
          if ((rc = apply_operator<double2>(c.p, X, AS, L, true, c.s))) return rc;
          double rt2;
          if ((rc = axpy_norm_host<V>(c, W, Bj, zc(-1.0), AS, &rt2))) return rc;
          stop = !(std::sqrt(rt2) / bnorm <= 2.0 * err);
        }
        if (stop) {
          c.conv32 = false;
          break;
        }
      }
    }
    if ((rc = op<V>(c, S, AS))) return rc;
    if ((rc = gram_to_host(c, S, AS, false, is_f32<V>(), StAS))) return rc;
    std::vector<zc> LU = StAS;
    if (!la::lu_factor(LU, piv, L)) return singular("block-COCG-rq (S^T A S)");
    alpha = QtQ;
    la::lu_solve(LU, piv, L, alpha, L);
    la::matmul(alpha, false, xi, ax, L);
    // X = X + S (alpha xi)
    if ((rc = update<V>(c, X, X, 1.0, S, nullptr, 0.0, ax, 0, nullptr, 0.0))) return rc;
    // W = Q - AS alpha
    negA.resize(alpha.size());
    for (size_t i = 0; i < alpha.size(); ++i) negA[i] = -alpha[i];
    // [Qn, tau] = qr(W); Qn written into AS (free until the next product)
    V* Qn = AS;
    bool qnqn_pending = false;
    if constexpr (std::is_same<V, double2>::value) {
      if (fused_gram<V>(c)) {
        std::vector<zc> G1;
        if ((rc = update_gram_launch(c, W, Q, 1.0, AS, negA, 1, true, 0)) || (rc = host_wait(c))) return rc;
        gram_unpack(c, 0, G1);
        if ((rc = cholqr2_fused(c, W, G1, Q1, Qn, tau, &qnqn_pending))) return rc;
      }
    }
    if (!fused_gram<V>(c)) {
      if ((rc = update<V>(c, W, Q, 1.0, AS, nullptr, 0.0, negA, 1, nullptr, 0.0))) return rc;
      if ((rc = cholqr2<V>(c, W, Q1, Qn, tau))) return rc;
    }
    std::vector<zc> xin;
    la::matmul(tau, false, xi, xin, L);
    xi.swap(xin);
    err = la::frob(xi) / bnorm;
    if (c.hist) c.hist->push_back(err);
    it = k;
    if (err < tol) break;
    if (qnqn_pending) {
      if ((rc = host_wait(c))) return rc;
      gram_unpack(c, 1, QntQn);
    } else if ((rc = gram_to_host(c, Qn, Qn, false, is_f32<V>(), QntQn))) {
      return rc;
    }
    la::matmul(tau, true, QntQn, rhs, L);  // tau^T (Qn^T Qn)
    LU = QtQ;
    if (!la::lu_factor(LU, piv, L)) return singular("block-COCG-rq (Q^T Q)");
    beta = rhs;
    la::lu_solve(LU, piv, L, beta, L);
    // Sn = Qn + S beta ; Q <- Qn by exchanging the Q and AS buffers (the
    // old Q is not read again; the next product overwrites AS)
    if ((rc = update<V>(c, Sn, Qn, 1.0, S, nullptr, 0.0, beta, 2, nullptr, 0.0))) return rc;
    std::swap(S, Sn);
    std::swap(Q, AS);
    QtQ.swap(QntQn);
  }
  *it_out = it;
  *err_out = err;
  return 0;
}

// ------------------------------------------- solvers with device L x L algebra
// For L <= kDlaMax the recurrence coefficients are computed on the device
// (kernels_dla.hip): Gram matrices feed the L x L solves directly, the
// block updates read their coefficients (and block-BiCGSTAB's scalar xi)
// from device memory, and an iteration waits on the host once, for its
// residual and a status word (singular system, failed or shifted Cholesky).
// Same recurrences, same step order and stopping rule as the host-algebra
// solvers above; the L x L arithmetic differs from the host's only in
// rounding.  Measured on the MI355X box (64^3 workload, L = 64: 38.9 against
// 38.0 ms per float64 solve; the 32-particle batch at L = 16: equal within
// 2%): the single-workgroup L x L kernels are latency-bound and take about
// as long as the host's AVX-512 routines plus the waits they save, so the
// host algebra stays the default and this path is selected with
// bdda_set_la_path(1) (one wait per iteration; a host with a slower CPU or
// a caller that wants no host work inside the iteration).
std::atomic<int> g_la_path{0};  // 0: host algebra, 1: device algebra when L <= kDlaMax

enum DlaSlot {
  D_QTQ, D_STAS, D_ALPHA, D_NALPHA, D_AX, D_G, D_R1, D_R1I, D_R2, D_R2I, D_R3, D_R3I, D_TAU, D_TAU2, D_XI, D_XIP,
  D_QNQN, D_RHS, D_BETA, D_LU, D_RV, D_RR, D_RZ, kDlaSlots
};
struct DlaMisc {  // layout of plan->dla_misc
  double err;
  int status;
  int pad;
  double2 sc[2];  // -xi, xi
  int piv[kDlaMax];
};

struct Dla {
  double2* slot[kDlaSlots];
  DlaMisc* misc;
  DlaMisc* host;
};

static bool use_dla(const Ctx& c) { return g_la_path.load() == 1 && c.L <= kDlaMax; }

static int dla_setup(Ctx& c, Dla& d) {
  bdda_plan* p = c.p;
  const size_t one = (size_t)c.L * c.L;
  int rc = p->dla.ensure(kDlaSlots * one * sizeof(double2));
  if (!rc) rc = p->dla_misc.ensure(sizeof(DlaMisc));
  if (!rc) rc = p->h_sync.ensure(sizeof(DlaMisc));
  if (!rc) rc = p->gram_part.ensure((size_t)gram_parts(c.L, c.rows) * ((c.L + 63) / 64) * ((c.L + 63) / 64) * 4096 *
                                    sizeof(double2));
  if (!rc) rc = p->dots_d.ensure(dot_parts() * sizeof(double2));
  if (!rc) rc = p->dots_n.ensure(dot_parts() * sizeof(double));
  if (rc) return rc;
  for (int i = 0; i < kDlaSlots; ++i) d.slot[i] = p->dla.as<double2>() + i * one;
  d.misc = p->dla_misc.as<DlaMisc>();
  d.host = p->h_sync.as<DlaMisc>();
  BDDA_HIP(hipMemsetAsync(&d.misc->status, 0, sizeof(int), c.s));
  return 0;
}

static int gram_dev(Ctx& c, const double2* A, const double2* B, bool conj, double2* out) {
  launch_gram<double2>(A, B, c.rows, c.L, conj, c.p->gram_part.as<double2>(), gram_parts(c.L, c.rows), out, c.s);
  BDDA_HIP(hipGetLastError());
  return 0;
}

// The iteration's one wait: err and status to the host.
static int dla_sync(Ctx& c, Dla& d, double* err, int* status) {
  BDDA_HIP(hipMemcpyAsync(d.host, d.misc, 2 * sizeof(double), hipMemcpyDeviceToHost, c.s));
  BDDA_HIP(hipStreamSynchronize(c.s));
  *err = d.host->err;
  *status = d.host->status;
  return 0;
}

// Y = a X + W1 C with C on the device (coefficient slot).
static int update_dev(Ctx& c, double2* Y, const double2* X, zc a, const double2* W1, const double2* W2,
                      const double2* C, const double2* E, const double2* sc) {
  launch_update<double2>(Y, X, make_double2(a.real(), a.imag()), W1, W2, make_double2(0.0, 0.0), C, E,
                         make_double2(0.0, 0.0), c.rows, c.L, c.s, sc);
  BDDA_HIP(hipGetLastError());
  return 0;
}

// Y = a X + W1 C (C on the device) on the matrix cores with the Gram matrix
// of Y (Y^H Y when conj, else Y^T Y) into gram_out (device); complex128
// blocks with L <= kUpdateGramMaxL.
static int update_gram_dev(Ctx& c, double2* Y, const double2* X, zc a, const double2* W1, const double2* C,
                           bool conj, double2* gram_out) {
  if (int rc = c.p->gram_part.ensure((size_t)update_gram_parts(c.rows) * 32 * 32 * sizeof(double2))) return rc;
  launch_update_gram(Y, X, make_double2(a.real(), a.imag()), W1, nullptr, make_double2(0.0, 0.0), C, nullptr,
                     make_double2(0.0, 0.0), c.rows, c.L, conj ? 1 : 2, c.p->gram_part.as<double2>(), gram_out, c.s);
  BDDA_HIP(hipGetLastError());
  return 0;
}

// Cholesky QR2 of W into Qn with tau = R2 R1 (device): the first pass
// shifts its Gram matrix if the plain factorisation fails (status
// DLA_SHIFTED; the caller then runs dla_qr_third_pass).  With g_ready, W^H W
// is already in slot D_G (from the update that wrote W).  Where the updates
// carry their Gram matrices (fused_gram), Q1^H Q1 comes from the update that
// writes Q1, and with qnqn the update that writes Qn leaves Qn^T Qn there.
static int cholqr2_dev(Ctx& c, Dla& d, const double2* W, double2* Q1, double2* Qn, bool g_ready = false,
                       double2* qnqn = nullptr) {
  int* st = &d.misc->status;
  const bool fused = fused_gram<double2>(c);
  if (!g_ready)
    if (int rc = gram_dev(c, W, W, true, d.slot[D_G])) return rc;
  dla_chol_inv(d.slot[D_G], c.L, (double)c.rows, true, d.slot[D_R1], d.slot[D_R1I], st, c.s);
  if (fused) {
    if (int rc = update_gram_dev(c, Q1, nullptr, 0.0, W, d.slot[D_R1I], true, d.slot[D_G])) return rc;
  } else {
    if (int rc = update_dev(c, Q1, nullptr, 0.0, W, nullptr, d.slot[D_R1I], nullptr, nullptr)) return rc;
    if (int rc = gram_dev(c, Q1, Q1, true, d.slot[D_G])) return rc;
  }
  dla_chol_inv(d.slot[D_G], c.L, (double)c.rows, false, d.slot[D_R2], d.slot[D_R2I], st, c.s);
  if (fused && qnqn) {
    if (int rc = update_gram_dev(c, Qn, nullptr, 0.0, Q1, d.slot[D_R2I], false, qnqn)) return rc;
  } else {
    if (int rc = update_dev(c, Qn, nullptr, 0.0, Q1, nullptr, d.slot[D_R2I], nullptr, nullptr)) return rc;
  }
  dla_matmul(d.slot[D_R2], false, d.slot[D_R1], d.slot[D_TAU], c.L, c.s);
  BDDA_HIP(hipGetLastError());
  return 0;
}

// Shifted CholeskyQR3's third pass on Qn (after a shifted first pass):
// Qn <- Qn R3^-1 (through Q1), tau <- R3 tau.
static int dla_qr_third_pass(Ctx& c, Dla& d, double2* Q1, double2* Qn) {
  int* st = &d.misc->status;
  if (int rc = gram_dev(c, Qn, Qn, true, d.slot[D_G])) return rc;
  dla_chol_inv(d.slot[D_G], c.L, (double)c.rows, false, d.slot[D_R3], d.slot[D_R3I], st, c.s);
  if (int rc = update_dev(c, Q1, nullptr, 0.0, Qn, nullptr, d.slot[D_R3I], nullptr, nullptr)) return rc;
  BDDA_HIP(hipMemcpyAsync(Qn, Q1, (size_t)c.rows * c.L * sizeof(double2), hipMemcpyDeviceToDevice, c.s));
  dla_matmul(d.slot[D_R3], false, d.slot[D_TAU], d.slot[D_TAU2], c.L, c.s);
  BDDA_HIP(hipMemcpyAsync(d.slot[D_TAU], d.slot[D_TAU2], (size_t)c.L * c.L * sizeof(double2),
                          hipMemcpyDeviceToDevice, c.s));
  return 0;
}

// The Householder last resort on the device-algebra path: Qn from W on the
// host, tau into its slot, the status cleared (called only without
// DLA_SINGULAR, which must survive).
static int householder_dev(Ctx& c, Dla& d, const double2* W, double2* Qn) {
  std::vector<zc> tau;
  if (int rc = householder_host<double2>(c, W, Qn, tau)) return rc;
  BDDA_HIP(hipMemcpyAsync(d.slot[D_TAU], tau.data(), tau.size() * sizeof(zc), hipMemcpyHostToDevice, c.s));
  BDDA_HIP(hipMemsetAsync(&d.misc->status, 0, sizeof(int), c.s));
  BDDA_HIP(hipStreamSynchronize(c.s));
  return 0;
}

static int dla_status_error(int status, const char* lu_what) {
  if (status & DLA_SINGULAR) return singular(lu_what);
  if (status & DLA_NOT_PD) return singular("block QR (Cholesky of the Gram matrix)");
  return 0;
}

// The mixed-precision check at the top of iteration k (as in the host-
// algebra solvers): true when the complex64 cycle must stop.
static int conv32_check(Ctx& c, const double2* X, const double2* Bj, double2* AS, double2* W, double bnorm,
                        double err, double tol, int k, bool* stop) {
  *stop = false;
  if (!(c.conv32 && c.check_every > 0 && k > 0 && k % c.check_every == 0)) return 0;
  *stop = !conv32_on_track(err, tol, k, c.check_every);
  if (!*stop) {
    if (int rc = apply_operator<double2>(c.p, X, AS, c.L, true, c.s)) return rc;
    double rt2;
    if (int rc = axpy_norm_host<double2>(c, W, Bj, zc(-1.0), AS, &rt2)) return rc;
    *stop = !(std::sqrt(rt2) / bnorm <= 2.0 * err);
  }
  if (*stop) c.conv32 = false;
  return 0;
}

// Block-COCG-rq (bl_krylov.py:58-118) with device L x L algebra.
static int solve_cocg_dev(Ctx& c, const double2* Bj, double2* X, bool x_zero, double bnorm, double tol,
                          int itermax, double2** ws, int* it_out, double* err_out) {
  const size_t bytes = (size_t)c.rows * c.L * sizeof(double2);
  const size_t lbytes = (size_t)c.L * c.L * sizeof(double2);
  double2 *Q = ws[0], *S = ws[1], *AS = ws[2], *W = ws[3], *Q1 = ws[4], *Sn = ws[5], *R0 = ws[6];
  Dla d;
  int rc = dla_setup(c, d);
  if (rc) return rc;
  if (x_zero) {
    BDDA_HIP(hipMemsetAsync(X, 0, bytes, c.s));
    BDDA_HIP(hipMemcpyAsync(R0, Bj, bytes, hipMemcpyDeviceToDevice, c.s));
  } else {
    if ((rc = op<double2>(c, X, AS))) return rc;
    launch_axpy_norm<double2>(R0, Bj, make_double2(-1.0, 0.0), AS, c.rows * c.L, nullptr, dot_parts(), c.s);
    BDDA_HIP(hipGetLastError());
  }
  // [Q, xi] = qr(R0), S = Q, QtQ = Q^T Q
  if ((rc = cholqr2_dev(c, d, R0, Q1, Q))) return rc;
  double e0;
  int st;
  if ((rc = dla_sync(c, d, &e0, &st))) return rc;
  if (st & DLA_SHIFTED) {
    int st3;
    BDDA_HIP(hipMemsetAsync(&d.misc->status, 0, sizeof(int), c.s));
    if ((rc = dla_qr_third_pass(c, d, Q1, Q)) || (rc = dla_sync(c, d, &e0, &st3))) return rc;
    st = (st & ~DLA_SHIFTED) | st3;
  }
  if ((st & DLA_NOT_PD) && !(st & DLA_SINGULAR)) {
    if ((rc = householder_dev(c, d, R0, Q))) return rc;
    st = 0;
  }
  if ((rc = dla_status_error(st, "block QR (Cholesky of W^H W)"))) return rc;
  BDDA_HIP(hipMemcpyAsync(d.slot[D_XI], d.slot[D_TAU], lbytes, hipMemcpyDeviceToDevice, c.s));
  BDDA_HIP(hipMemcpyAsync(S, Q, bytes, hipMemcpyDeviceToDevice, c.s));
  if ((rc = gram_dev(c, Q, Q, false, d.slot[D_QTQ]))) return rc;
  int it = 0;
  double err = INFINITY;
  for (int k = 0; k < itermax; ++k) {
    bool stop = false;
    if ((rc = conv32_check(c, X, Bj, AS, W, bnorm, err, tol, k, &stop))) return rc;
    if (stop) break;
    if ((rc = op<double2>(c, S, AS))) return rc;
    if ((rc = gram_dev(c, S, AS, false, d.slot[D_STAS]))) return rc;
    dla_lu_solve(d.slot[D_STAS], d.slot[D_QTQ], c.L, c.L, false, d.slot[D_ALPHA], d.slot[D_NALPHA], nullptr,
                 nullptr, &d.misc->status, c.s);
    dla_matmul(d.slot[D_ALPHA], false, d.slot[D_XI], d.slot[D_AX], c.L, c.s);
    // X = X + S (alpha xi) ; W = Q - AS alpha
    if ((rc = update_dev(c, X, X, 1.0, S, nullptr, d.slot[D_AX], nullptr, nullptr))) return rc;
    // [Qn, tau] = qr(W) into AS (free until the next product); xi = tau xi.
    // With fused Gram matrices W^H W comes with W and Qn^T Qn with Qn (the
    // latter stands unless a fallback below rewrites Qn)
    const bool fused = fused_gram<double2>(c);
    bool qnqn_ready = fused;
    if (fused) {
      if ((rc = update_gram_dev(c, W, Q, 1.0, AS, d.slot[D_NALPHA], true, d.slot[D_G]))) return rc;
    } else if ((rc = update_dev(c, W, Q, 1.0, AS, nullptr, d.slot[D_NALPHA], nullptr, nullptr))) {
      return rc;
    }
    double2* Qn = AS;
    if ((rc = cholqr2_dev(c, d, W, Q1, Qn, fused, fused ? d.slot[D_QNQN] : nullptr))) return rc;
    BDDA_HIP(hipMemcpyAsync(d.slot[D_XIP], d.slot[D_XI], lbytes, hipMemcpyDeviceToDevice, c.s));
    dla_matmul(d.slot[D_TAU], false, d.slot[D_XIP], d.slot[D_XI], c.L, c.s);
    dla_frob(d.slot[D_XI], c.L, bnorm, &d.misc->err, c.s);
    if ((rc = dla_sync(c, d, &err, &st))) return rc;
    if ((st & DLA_SHIFTED) && !(st & (DLA_SINGULAR | DLA_NOT_PD))) {
      int st3;
      BDDA_HIP(hipMemsetAsync(&d.misc->status, 0, sizeof(int), c.s));
      if ((rc = dla_qr_third_pass(c, d, Q1, Qn))) return rc;
      qnqn_ready = false;
      dla_matmul(d.slot[D_TAU], false, d.slot[D_XIP], d.slot[D_XI], c.L, c.s);
      dla_frob(d.slot[D_XI], c.L, bnorm, &d.misc->err, c.s);
      if ((rc = dla_sync(c, d, &err, &st3))) return rc;
      st = (st & ~DLA_SHIFTED) | st3;
    }
    if ((st & DLA_NOT_PD) && !(st & DLA_SINGULAR)) {
      qnqn_ready = false;
      if ((rc = householder_dev(c, d, W, Qn))) return rc;
      dla_matmul(d.slot[D_TAU], false, d.slot[D_XIP], d.slot[D_XI], c.L, c.s);
      dla_frob(d.slot[D_XI], c.L, bnorm, &d.misc->err, c.s);
      if ((rc = dla_sync(c, d, &err, &st))) return rc;
    }
    if ((rc = dla_status_error(st, "block-COCG-rq (S^T A S)"))) return rc;
    if (c.hist) c.hist->push_back(err);
    it = k;
    if (err < tol) break;
    // beta = (Q^T Q)^-1 tau^T (Qn^T Qn) ; Sn = Qn + S beta ; Q <- Qn
    if (!qnqn_ready && (rc = gram_dev(c, Qn, Qn, false, d.slot[D_QNQN]))) return rc;
    dla_matmul(d.slot[D_TAU], true, d.slot[D_QNQN], d.slot[D_RHS], c.L, c.s);
    dla_lu_solve(d.slot[D_QTQ], d.slot[D_RHS], c.L, c.L, false, d.slot[D_BETA], nullptr, nullptr, nullptr,
                 &d.misc->status, c.s);
    if ((rc = update_dev(c, Sn, Qn, 1.0, S, nullptr, d.slot[D_BETA], nullptr, nullptr))) return rc;
    std::swap(S, Sn);
    BDDA_HIP(hipMemcpyAsync(Q, Qn, bytes, hipMemcpyDeviceToDevice, c.s));
    BDDA_HIP(hipMemcpyAsync(d.slot[D_QTQ], d.slot[D_QNQN], lbytes, hipMemcpyDeviceToDevice, c.s));
  }
  *it_out = it;
  *err_out = err;
  return 0;
}

// Block-BiCGSTAB (bl_krylov.py:121-187) with device L x L algebra.
static int solve_bicgstab_dev(Ctx& c, const double2* Bj, double2* X, bool x_zero, double bnorm, double tol,
                              int itermax, double2** ws, int* it_out, double* err_out) {
  const long long n = c.rows * c.L;
  const size_t bytes = n * sizeof(double2);
  double2 *R = ws[0], *P = ws[1], *R0 = ws[2], *Vb = ws[3], *T = ws[4], *Z = ws[5], *Pn = ws[6];
  Dla d;
  int rc = dla_setup(c, d);
  if (rc) return rc;
  if (x_zero) {
    BDDA_HIP(hipMemsetAsync(X, 0, bytes, c.s));
    BDDA_HIP(hipMemcpyAsync(R, Bj, bytes, hipMemcpyDeviceToDevice, c.s));
  } else {
    if ((rc = op<double2>(c, X, T))) return rc;
    launch_axpy_norm<double2>(R, Bj, make_double2(-1.0, 0.0), T, n, nullptr, dot_parts(), c.s);
    BDDA_HIP(hipGetLastError());
  }
  BDDA_HIP(hipMemcpyAsync(P, R, bytes, hipMemcpyDeviceToDevice, c.s));
  BDDA_HIP(hipMemcpyAsync(R0, R, bytes, hipMemcpyDeviceToDevice, c.s));
  const double2* sc = d.misc->sc;  // -xi, xi
  int it = 0, st = 0;
  double err = INFINITY;
  for (int k = 0; k < itermax; ++k) {
    bool stop = false;
    if ((rc = conv32_check(c, X, Bj, T, Z, bnorm, err, tol, k, &stop))) return rc;
    if (stop) break;
    if ((rc = op<double2>(c, P, Vb))) return rc;
    if ((rc = gram_dev(c, R0, Vb, true, d.slot[D_RV])) || (rc = gram_dev(c, R0, R, true, d.slot[D_RR]))) return rc;
    dla_lu_solve(d.slot[D_RV], d.slot[D_RR], c.L, c.L, false, d.slot[D_ALPHA], d.slot[D_NALPHA], d.slot[D_LU],
                 d.misc->piv, &d.misc->status, c.s);
    // T = R - V alpha
    if ((rc = update_dev(c, T, R, 1.0, Vb, nullptr, d.slot[D_NALPHA], nullptr, nullptr))) return rc;
    if ((rc = op<double2>(c, T, Z))) return rc;
    // xi = tr(Z^H T) / tr(Z^H Z) from one pass over T and Z
    launch_dot<double2>(T, Z, n, c.p->dots_d.as<double2>(), c.p->dots_n.as<double>(), dot_parts(), c.s);
    BDDA_HIP(hipGetLastError());
    dla_xi(c.p->dots_d.as<double2>(), c.p->dots_n.as<double>(), dot_parts(), &d.misc->status, d.misc->sc, c.s);
    // X = X + P alpha + xi T ; R = T - xi Z with its norm ; R0^H Z
    if ((rc = update_dev(c, X, X, 1.0, P, nullptr, d.slot[D_ALPHA], T, sc))) return rc;
    launch_axpy_norm<double2>(R, T, make_double2(0.0, 0.0), Z, n, c.p->dots_n.as<double>(), dot_parts(), c.s, sc);
    BDDA_HIP(hipGetLastError());
    if ((rc = gram_dev(c, R0, Z, true, d.slot[D_RZ]))) return rc;
    dla_norm_err(c.p->dots_n.as<double>(), dot_parts(), bnorm, &d.misc->err, c.s);
    if ((rc = dla_sync(c, d, &err, &st))) return rc;
    if ((rc = dla_status_error(st, "block-BiCGSTAB (R0^H V)"))) return rc;
    if (c.hist) c.hist->push_back(err);
    it = k;
    if (err < tol) break;
    // beta = (R0^H V)^-1 (-R0^H Z) ; Pn = R + (P - xi V) beta
    dla_lu_resolve(d.slot[D_LU], d.misc->piv, d.slot[D_RZ], c.L, c.L, true, d.slot[D_BETA], &d.misc->status, c.s);
    if ((rc = update_dev(c, Pn, R, 1.0, P, Vb, d.slot[D_BETA], nullptr, sc))) return rc;
    std::swap(P, Pn);
  }
  *it_out = it;
  *err_out = err;
  return 0;
}

template <typename V>
static int run_solver(Ctx& c, int solver, const V* Bj, V* X, bool x_zero, double bnorm, double tol, int itermax,
                      V** ws, int* it, double* err) {
  if constexpr (std::is_same<V, double2>::value) {
    if (use_dla(c)) {
      if (solver == BDDA_SOLVER_BICGSTAB) return solve_bicgstab_dev(c, Bj, X, x_zero, bnorm, tol, itermax, ws, it, err);
      return solve_cocg_dev(c, Bj, X, x_zero, bnorm, tol, itermax, ws, it, err);
    }
  }
  if (solver == BDDA_SOLVER_BICGSTAB) return solve_bicgstab<V>(c, Bj, X, x_zero, bnorm, tol, itermax, ws, it, err);
  return solve_cocg<V>(c, Bj, X, x_zero, bnorm, tol, itermax, ws, it, err);
}

static int frob2(Ctx& c, const double2* A, double* out) { return dots_to_host<double2>(c, A, A, nullptr, out); }

}  // namespace bdda

using namespace bdda;

// ------------------------------------------- allocation-free observables
// Workgroup partial sums of bdda_observables_noalloc live in this
// file-scope device array (one copy per GPU, held until the process ends)
// instead of plan memory allocated on demand.  Host code never uses the
// array's name as an address: each device's copy is looked up once with
// hipGetSymbolAddress and the pointer kept per device.
namespace {
constexpr int kNoAllocMaxL = 256;
constexpr size_t kObsScratchElems = 256 * (size_t)kNoAllocMaxL * 5;  // obs_parts() x L x 5
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {nullptr};
}  // namespace
static __device__ __align__(256) double2 scratch[kObsScratchElems];

// The calling thread's current device's copy of `scratch` (looked up on
// the first call on that device).
static int scratch_ptr(void** out) {
  int dev = 0;
  BDDA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("allocation-free observables: device index out of range");
    return BDDA_E_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return BDDA_E_HIP;
    }
    g_scratch_dev[dev] = ptr;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

// ====================================================================== C ABI
extern "C" {

const char* bdda_last_error(void) { return g_err.c_str(); }
int bdda_version(void) { return 100; }

int bdda_set_la_path(int path) {
  if (path < 0 || path > 1) {
    set_error("bdda_set_la_path: path must be 0 (host algebra) or 1 (device algebra for L <= 64)");
    return BDDA_E_ARG;
  }
  g_la_path = path;
  return 0;
}

int bdda_set_operator_groups(int groups) {
  if (groups < 1 || groups > bdda_plan::kMaxGroups) {
    set_error("bdda_set_operator_groups: groups must be 1.." + std::to_string(bdda_plan::kMaxGroups));
    return BDDA_E_ARG;
  }
  g_op_groups.store(groups);
  return 0;
}

int bdda_set_fft_path(int path) {
  if (path < 0 || path > 4) {
    set_error("bdda_set_fft_path: path must be 0 (auto), 1 (runtime-radix only), 2 (no dedicated x pass), 3 "
              "(the split-spectrum x pass on every eligible length) or 4 (the one-pass x pass everywhere)");
    return BDDA_E_ARG;
  }
  set_fft_path(path);
  return 0;
}

int bdda_plan_create(int nx, int ny, int nz, bdda_plan** out) {
  g_err.clear();
  if (!out || nx < 1 || ny < 1 || nz < 1) {
    set_error("bdda_plan_create: lattice dimensions must be >= 1");
    return BDDA_E_ARG;
  }
  std::unique_ptr<bdda_plan> p(new bdda_plan());
  const int n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    p->ax[a].n = n[a];
    // the reference's 2N grid when it is 5-smooth, else a length >= 2N - 1
    // (still an exact linear convolution)
    int m = conv_length(n[a], a);
    if (m < 2) m = 2;
    if (!factor_length(m, p->ax[a])) {
      set_error("bdda_plan_create: cannot factor FFT length");
      return BDDA_E_ARG;
    }
    if ((long long)m * 3 > 4096) {
      set_error("bdda_plan_create: axis too long (FFT length must be <= 1365)");
      return BDDA_E_ARG;
    }
    // twiddles exp(-2 pi i k / m)
    std::vector<double2> t64(m);
    std::vector<float2> t32(m);
    for (int k = 0; k < m; ++k) {
      const double ang = -2.0 * M_PI * (double)k / (double)m;
      t64[k] = make_double2(std::cos(ang), std::sin(ang));
      t32[k] = make_float2((float)t64[k].x, (float)t64[k].y);
    }
    int rc = p->tw64[a].ensure(m * sizeof(double2));
    if (!rc) rc = p->tw32[a].ensure(m * sizeof(float2));
    if (rc) return rc;
    BDDA_HIP(hipMemcpy(p->tw64[a].p, t64.data(), m * sizeof(double2), hipMemcpyHostToDevice));
    BDDA_HIP(hipMemcpy(p->tw32[a].p, t32.data(), m * sizeof(float2), hipMemcpyHostToDevice));
  }
  p->ncell = (long long)nx * ny * nz;
  *out = p.release();
  return 0;
}

int bdda_plan_destroy(bdda_plan* plan) {
  if (plan) {
    (void)hipDeviceSynchronize();
    delete plan;
  }
  return 0;
}

int bdda_plan_fft_dims(const bdda_plan* plan, int* m_out) {
  if (int rc = check_plan(plan)) return rc;
  for (int a = 0; a < 3; ++a) m_out[a] = plan->ax[a].m;
  return 0;
}

int bdda_plan_column_tile(const bdda_plan* p, int L) {
  if (check_plan(p) || L < 1) return -BDDA_E_ARG;
  return choose_lb(p, L);
}

int bdda_green_init(bdda_plan* p, double lf, double k) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  const long long M = (long long)p->ax[0].m * p->ax[1].m * p->ax[2].m;
  const long long H = (long long)p->ax[0].h * p->ax[1].h * p->ax[2].h;
  DevBuf full;
  // fresh buffers: a spectrum shared with other plans is never overwritten
  auto g64 = std::make_shared<DevBuf>(), g32 = std::make_shared<DevBuf>();
  int rc = full.ensure((size_t)M * 6 * sizeof(double2));
  if (!rc) rc = g64->ensure((size_t)H * 6 * sizeof(double2));
  if (!rc) rc = g32->ensure((size_t)H * 6 * sizeof(float2));
  if (rc) return rc;
  p->green_ready = false;
  hipStream_t s = 0;
  launch_green_build<double2>(full.as<double2>(), p->ax, lf, k, s);
  BDDA_HIP(hipGetLastError());
  const int mx = p->ax[0].m, my = p->ax[1].m, mz = p->ax[2].m;
  // in-place 3-D FFT of the 6 components, layout [x][y][z][6]
  for (int a = 2; a >= 0; --a) {
    LinePass lp;
    std::memset(&lp, 0, sizeof(lp));
    lp.mode = PASS_PLAIN;
    lp.m = p->ax[a].m;
    lp.nin = lp.m;
    lp.nout = lp.m;
    lp.nrad = p->ax[a].nrad;
    for (int i = 0; i < kMaxRadix; ++i) lp.rad[i] = p->ax[a].rad[i];
    lp.tw = p->tw64[a].p;
    lp.Q = 6;
    lp.nlo = 1;
    lp.src = full.p;
    lp.dst = full.p;
    if (a == 2) {  // z lines, points (x, y)
      lp.npts = mx * my;
      lp.ptb = lp.npts;
      lp.sb = lp.db = (long long)mz * 6;
      lp.se = lp.de = 6;
      lp.order_pe = 0;
    } else if (a == 1) {  // y lines, points (x, z)
      lp.npts = mx * mz;
      lp.ptb = mz;
      lp.sa = lp.da = (long long)my * mz * 6;
      lp.sb = lp.db = 6;
      lp.se = lp.de = (long long)mz * 6;
      lp.order_pe = 1;
    } else {  // x lines, points (y, z)
      lp.npts = my * mz;
      lp.ptb = lp.npts;
      lp.sb = lp.db = 6;
      lp.se = lp.de = (long long)my * mz * 6;
      lp.order_pe = 1;
    }
    launch_line_pass<double2>(lp, s);
    BDDA_CHECK_LAUNCH();
  }
  launch_green_halve<double2>(full.as<double2>(), g64->as<double2>(), p->ax, s);
  launch_green_halve<float2>(full.as<double2>(), g32->as<float2>(), p->ax, s);
  BDDA_HIP(hipGetLastError());
  BDDA_HIP(hipDeviceSynchronize());
  p->g64 = std::move(g64);
  p->g32 = std::move(g32);
  p->green_lf = lf;
  p->green_k = k;
  p->green_ready = true;
  return 0;
}

int bdda_green_share(bdda_plan* dst, const bdda_plan* src) {
  if (int rc = check_plan(dst)) return rc;
  if (int rc = check_plan(src)) return rc;
  g_err.clear();
  if (!src->green_ready) {
    set_error("bdda_green_share: the source plan has no Green spectrum (call bdda_green_init on it)");
    return BDDA_E_STATE;
  }
  for (int a = 0; a < 3; ++a)
    if (dst->ax[a].n != src->ax[a].n || dst->ax[a].m != src->ax[a].m) {
      set_error("bdda_green_share: the plans cover different cuboids");
      return BDDA_E_ARG;
    }
  if (dst == src) return 0;
  dst->g64 = src->g64;
  dst->g32 = src->g32;
  dst->green_lf = src->green_lf;
  dst->green_k = src->green_k;
  dst->green_ready = true;
  return 0;
}

int bdda_green_download(const bdda_plan* p, void* out_host) {
  if (int rc = check_plan(p)) return rc;
  if (!p->green_ready) {
    set_error("bdda_green_download: call bdda_green_init first");
    return BDDA_E_STATE;
  }
  const long long M = (long long)p->ax[0].m * p->ax[1].m * p->ax[2].m;
  DevBuf full;
  if (int rc = full.ensure((size_t)M * 9 * sizeof(double2))) return rc;
  launch_green_expand(p->g64->as<double2>(), full.as<double2>(), p->ax, nullptr, 0);
  BDDA_HIP(hipGetLastError());
  BDDA_HIP(hipMemcpy(out_host, full.p, (size_t)M * 9 * sizeof(double2), hipMemcpyDeviceToHost));
  return 0;
}

int bdda_green_tensor(const bdda_plan* p, double lf, double k, void* out_host) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  const long long M = (long long)p->ax[0].m * p->ax[1].m * p->ax[2].m;
  DevBuf g6, g9;
  int rc = g6.ensure((size_t)M * 6 * sizeof(double2));
  if (!rc) rc = g9.ensure((size_t)M * 9 * sizeof(double2));
  if (rc) return rc;
  launch_green_build<double2>(g6.as<double2>(), p->ax, lf, k, 0);
  launch_green_sym9(g6.as<double2>(), g9.as<double2>(), M, 0);
  BDDA_HIP(hipGetLastError());
  BDDA_HIP(hipMemcpy(out_host, g9.p, (size_t)M * 9 * sizeof(double2), hipMemcpyDeviceToHost));
  return 0;
}

int bdda_mvp_offdiag(bdda_plan* p, const void* P_dev, void* Y_dev, int L, int precision, void* stream) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  if (!p->green_ready) {
    set_error("bdda_mvp_offdiag: call bdda_green_init first");
    return BDDA_E_STATE;
  }
  if (L < 1 || !P_dev || !Y_dev) {
    set_error("bdda_mvp_offdiag: bad arguments");
    return BDDA_E_ARG;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (precision == BDDA_PREC_FP64)
    return apply_operator<double2>(p, (const double2*)P_dev, (double2*)Y_dev, L, false, s);
  // complex64 transforms on a converted copy
  const long long n = 3 * p->ncell * (long long)L;
  DevBuf a, b;
  int rc = a.ensure(n * sizeof(float2));
  if (!rc) rc = b.ensure(n * sizeof(float2));
  if (rc) return rc;
  launch_convert<float2, double2>(a.as<float2>(), (const double2*)P_dev, n, 1.0, s);
  if ((rc = apply_operator<float2>(p, a.as<float2>(), b.as<float2>(), L, false, s))) return rc;
  launch_convert<double2, float2>((double2*)Y_dev, b.as<float2>(), n, 1.0, s);
  BDDA_HIP(hipGetLastError());
  BDDA_HIP(hipStreamSynchronize(s));
  return 0;
}

int bdda_profile_enable(bdda_plan* p, int on) {
  if (int rc = check_plan(p)) return rc;
  p->profiling = on != 0;
  p->ev_used = 0;
  p->ev_pass.clear();
  return 0;
}

int bdda_profile_read(bdda_plan* p, double* ms_sum, int64_t* count) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  for (int i = 0; i < 10; ++i) {
    ms_sum[i] = 0.0;
    count[i] = 0;
  }
  BDDA_HIP(hipDeviceSynchronize());
  for (size_t i = 0; i < p->ev_pass.size(); ++i) {
    float ms = 0.f;
    BDDA_HIP(hipEventElapsedTime(&ms, p->ev_pool[2 * i], p->ev_pool[2 * i + 1]));
    ms_sum[p->ev_pass[i]] += ms;
    count[p->ev_pass[i]] += 1;
  }
  p->ev_used = 0;
  p->ev_pass.clear();
  return 0;
}

int bdda_set_target(bdda_plan* p, const int64_t* address, int64_t N, const double* pos, const void* diag,
                    const void* jpre) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  if (N < 1 || !address || !diag || !jpre) {
    set_error("bdda_set_target: bad arguments");
    return BDDA_E_ARG;
  }
  for (int64_t i = 0; i < N; ++i) {
    if (address[i] < 0 || address[i] >= p->ncell || (i > 0 && address[i] <= address[i - 1])) {
      set_error("bdda_set_target: addresses must be increasing and inside the cuboid");
      return BDDA_E_ARG;
    }
  }
  {
    const long long ny = p->ax[1].n, nz = p->ax[2].n;
    int ext[3] = {0, 0, 0};
    for (int64_t i = 0; i < N; ++i) {
      const long long a = address[i];
      ext[0] = std::max(ext[0], (int)(a / (ny * nz)) + 1);
      ext[1] = std::max(ext[1], (int)((a / nz) % ny) + 1);
      ext[2] = std::max(ext[2], (int)(a % nz) + 1);
    }
    for (int d = 0; d < 3; ++d) p->occ[d] = ext[d];
  }
  std::vector<int> lines;
  std::vector<unsigned char> mask((size_t)p->ax[0].n * p->ax[1].n, 0);
  for (int64_t i = 0; i < N; ++i) {
    const int l = (int)(address[i] / p->ax[2].n);  // addresses ascend, so lines do too
    if (lines.empty() || lines.back() != l) lines.push_back(l);
    mask[l] = 1;
  }
  int rc = p->zlines.ensure(lines.size() * sizeof(int));
  if (!rc) rc = p->lmask.ensure(mask.size());
  if (rc) return rc;
  BDDA_HIP(hipMemcpy(p->zlines.p, lines.data(), lines.size() * sizeof(int), hipMemcpyHostToDevice));
  BDDA_HIP(hipMemcpy(p->lmask.p, mask.data(), mask.size(), hipMemcpyHostToDevice));
  p->nzl = (int)lines.size();
  ++p->lines_version;
  rc = p->inv.ensure(p->ncell * sizeof(int));
  if (!rc) rc = p->pos.ensure(N * 3 * sizeof(double));
  if (!rc) rc = p->diag64.ensure(3 * N * sizeof(double2));
  if (!rc) rc = p->jpre64.ensure(3 * N * sizeof(double2));
  if (rc) return rc;
  DevBuf addr;
  if ((rc = addr.ensure(N * sizeof(long long)))) return rc;
  BDDA_HIP(hipMemcpy(addr.p, address, N * sizeof(long long), hipMemcpyHostToDevice));
  launch_fill_inv(p->inv.as<int>(), p->ncell, addr.as<long long>(), N, 0);
  if (pos) BDDA_HIP(hipMemcpy(p->pos.p, pos, N * 3 * sizeof(double), hipMemcpyHostToDevice));
  p->has_pos = pos != nullptr;
  BDDA_HIP(hipMemcpy(p->diag64.p, diag, 3 * N * sizeof(double2), hipMemcpyHostToDevice));
  BDDA_HIP(hipMemcpy(p->jpre64.p, jpre, 3 * N * sizeof(double2), hipMemcpyHostToDevice));
  BDDA_HIP(hipGetLastError());
  BDDA_HIP(hipDeviceSynchronize());
  p->N = N;
  p->target_ready = true;
  return 0;
}

int bdda_plan_set_column_tile(bdda_plan* p, int lb) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  if (lb < 0 || lb > lb_cap(p)) {
    set_error("bdda_plan_set_column_tile: tile must be 0 (automatic) or 1.." + std::to_string(lb_cap(p)));
    return BDDA_E_ARG;
  }
  p->lb_force = lb;
  return 0;
}

int bdda_plan_set_x_kernel(bdda_plan* p, int kernel) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  if (kernel < 0 || kernel > 3) {
    set_error("bdda_plan_set_x_kernel: kernel must be 0 (the process-wide rule), 1 (split-spectrum where "
              "eligible), 2 (one-pass) or 3 (split-spectrum where the radix pair is balanced)");
    return BDDA_E_ARG;
  }
  p->x_kernel = kernel;
  return 0;
}

int bdda_plan_x_kernel(const bdda_plan* p) {
  if (int rc = check_plan(p)) return -rc;
  return p->x_kernel;
}

// Times the occupied operator (one warm-up, then three applications between
// HIP events on the caller's stream) for the column tiles 4, 8 and the
// automatic choice, within the plan's cap, each with the split-spectrum and
// the one-pass x pass, and keeps the fastest pair.
int bdda_plan_autotune(bdda_plan* p, int L, int precision, int* lb_out, void* stream) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  if (!p->green_ready || !p->target_ready) {
    set_error("bdda_plan_autotune: call bdda_green_init and bdda_set_target first");
    return BDDA_E_STATE;
  }
  if (L < 1 || (precision != BDDA_PREC_FP64 && precision != BDDA_PREC_MIXED)) {
    set_error("bdda_plan_autotune: bad arguments");
    return BDDA_E_ARG;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int cap = lb_cap(p);
  p->lb_force = 0;
  std::vector<int> cand = {choose_lb(p, L)};
  for (int lb : {8, 4})
    if (lb <= cap && std::find(cand.begin(), cand.end(), lb) == cand.end()) cand.push_back(lb);
  DevBuf S, AS;
  const size_t bytes = (size_t)p->N * 3 * L * sizeof(double2);
  if (int rc = S.ensure(bytes)) return rc;
  if (int rc = AS.ensure(bytes)) return rc;
  BDDA_HIP(hipMemsetAsync(S.p, 0, bytes, s));
  hipEvent_t e0, e1;
  BDDA_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_error("bdda_plan_autotune: hipEventCreate failed");
    return BDDA_E_HIP;
  }
  int best = cand[0], best_xk = 0, rc = 0;
  float best_ms = INFINITY;
  for (int c = 0; c < (int)cand.size() * 2 && !rc; ++c) {
    const int lb = cand[c / 2], xk = 1 + c % 2;
    p->lb_force = lb;
    p->x_kernel = xk;
    auto apply = [&]() {
      return precision == BDDA_PREC_MIXED
                 ? apply_operator<float2, double2>(p, S.as<double2>(), AS.as<double2>(), L, true, s)
                 : apply_operator<double2>(p, S.as<double2>(), AS.as<double2>(), L, true, s);
    };
    if ((rc = apply())) break;
    if (hipEventRecord(e0, s) != hipSuccess) { rc = BDDA_E_HIP; break; }
    for (int r = 0; r < 3 && !rc; ++r) rc = apply();
    if (rc) break;
    float ms = 0.0f;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      set_error("bdda_plan_autotune: event timing failed");
      rc = BDDA_E_HIP;
      break;
    }
    if (ms < best_ms) {
      best_ms = ms;
      best = lb;
      best_xk = xk;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  p->lb_force = rc ? 0 : best;
  p->x_kernel = rc ? 0 : best_xk;
  if (!rc && lb_out) *lb_out = best;
  return rc;
}

int bdda_block_mvp(bdda_plan* p, const void* S_dev, void* AS_dev, int L, void* stream) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  if (!p->green_ready || !p->target_ready) {
    set_error("bdda_block_mvp: call bdda_green_init and bdda_set_target first");
    return BDDA_E_STATE;
  }
  if (L < 1 || !S_dev || !AS_dev) {
    set_error("bdda_block_mvp: bad arguments");
    return BDDA_E_ARG;
  }
  return apply_operator<double2>(p, (const double2*)S_dev, (double2*)AS_dev, L, true,
                                 reinterpret_cast<hipStream_t>(stream));
}

int bdda_solve(bdda_plan* p, int solver, int precision, const void* B_dev, const void* X0_dev, void* X_dev, int L,
               double tol, int itermax, int refine_every, int* iter_fin, double* err_fin, double* true_res,
               double* err_hist, int* flags_out, void* stream) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  if (!p->green_ready || !p->target_ready) {
    set_error("bdda_solve: call bdda_green_init and bdda_set_target first");
    return BDDA_E_STATE;
  }
  if (L < 1 || L > kMaxL || !B_dev || !X_dev || itermax < 0 ||
      (solver != BDDA_SOLVER_COCG_RQ && solver != BDDA_SOLVER_BICGSTAB) ||
      (precision != BDDA_PREC_FP64 && precision != BDDA_PREC_MIXED)) {
    set_error("bdda_solve: bad arguments (1 <= L <= 4096, known solver and precision)");
    return BDDA_E_ARG;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  std::vector<double> hist;
  Ctx c{p, L, 3 * p->N, s, &hist};
  const long long n = c.rows * L;
  const size_t b64 = n * sizeof(double2), b32 = n * sizeof(float2);
  int rc = p->Bj64.ensure(b64);
  for (int i = 0; i < 9 && !rc; ++i) rc = p->blk[i].ensure(b64);
  if (rc) return rc;
  double2* X = (double2*)X_dev;
  double2* Bj = p->Bj64.as<double2>();
  double2* tmp = p->blk[8].as<double2>();
  // Bj = jpre * B, bnorm = ||Bj||_F  (bl_krylov.py:87-88)
  launch_scale_rows<double2>(Bj, (const double2*)B_dev, p->jpre64.as<double2>(), c.rows, L, s);
  BDDA_HIP(hipGetLastError());
  double bn2 = 0.0;
  if ((rc = frob2(c, Bj, &bn2))) return rc;
  const double bnorm = std::sqrt(bn2);
  // warm start: keep X0 only if its preconditioned residual is below ||Bj||
  bool x_zero = true;
  int flags = 0;
  if (X0_dev) {
    if ((rc = apply_operator<double2>(p, (const double2*)X0_dev, tmp, L, true, s))) return rc;
    double r2 = 0.0;
    if ((rc = axpy_norm_host<double2>(c, tmp, Bj, zc(-1.0), tmp, &r2))) return rc;
    if (std::sqrt(r2) <= bnorm) {
      x_zero = false;
      flags |= 1;
      if ((const void*)X0_dev != (void*)X) BDDA_HIP(hipMemcpyAsync(X, X0_dev, b64, hipMemcpyDeviceToDevice, s));
    }
  }
  int it = 0;
  double err = INFINITY;
  if (precision == BDDA_PREC_FP64) {
    double2* ws[7];
    for (int i = 0; i < 7; ++i) ws[i] = p->blk[i].as<double2>();
    if ((rc = run_solver<double2>(c, solver, Bj, X, x_zero, bnorm, tol, itermax, ws, &it, &err))) return rc;
  } else {
    // mixed precision: the Goodman convolution (nearly all of the operator's
    // bytes) runs in complex64 between complex128 scatter / gather
    // boundaries; the Krylov blocks, their updates, Gram matrices and L x L
    // solves stay float64.  A cycle solves for the correction D on the
    // float64 true residual R = Bj - A X with the complex64 convolution;
    // every refine_every iterations it computes its true residual with the
    // complex128 operator, and once that exceeds twice the recursive one
    // (complex64 rounding amplified by the recurrence) it stops; it also
    // stops when its recursive residual is off a log-linear course to the
    // tolerance.  After a cycle X += D and R is recomputed; a cycle that made
    // R worse is undone.  Classic iterative refinement: while complex64
    // cycles reach their own tolerance and at least halve the true residual,
    // the next cycle restarts from the new true residual in complex64 again;
    // a cycle stopped by its check (a hard solve, where complex64 rounding
    // costs the recurrence its conjugacy) or one that does not halve the
    // residual hands the rest of the iteration budget to complex128.
    if ((rc = p->blk[9].ensure(b64)) || (rc = p->blk[10].ensure(b64))) return rc;
    double2* ws[7];
    for (int i = 0; i < 7; ++i) ws[i] = p->blk[i].as<double2>();
    double2* R64 = p->blk[7].as<double2>();
    double2* D = p->blk[9].as<double2>();
    double2* Rsave = p->blk[10].as<double2>();
    bool conv32 = true;
    int cycles = 0;
    if (x_zero) {
      BDDA_HIP(hipMemsetAsync(X, 0, b64, s));
      BDDA_HIP(hipMemcpyAsync(R64, Bj, b64, hipMemcpyDeviceToDevice, s));
    } else {
      if ((rc = apply_operator<double2>(p, X, tmp, L, true, s))) return rc;
      double d;
      if ((rc = axpy_norm_host<double2>(c, R64, Bj, zc(-1.0), tmp, &d))) return rc;
    }
    double r2;
    if ((rc = frob2(c, R64, &r2))) return rc;
    double res = std::sqrt(r2) / bnorm;
    int total = 0;
    err = res;
    while (res >= tol && total < itermax) {
      const double rs = std::sqrt(r2);
      // the complex64 convolution's 1e-7 rounding bounds what a complex64
      // cycle reaches
      double inner_tol = 0.5 * tol * bnorm / rs;
      if (conv32 && inner_tol < 2e-6) inner_tol = 2e-6;
      ++cycles;
      std::vector<double> inner;
      c.hist = &inner;
      c.conv32 = conv32;
      c.check_every = refine_every > 0 ? refine_every : 0;
      const bool was32 = conv32;
      int iit = 0;
      double ierr = 0.0;
      rc = run_solver<double2>(c, solver, R64, D, true, rs, inner_tol, itermax - total, ws, &iit, &ierr);
      conv32 = c.conv32;
      c.conv32 = false;
      c.check_every = 0;
      c.hist = &hist;
      // a singular block QR / L x L system ends the cycle; the solvers update
      // D in place, so the correction reached so far stands
      if (rc == BDDA_E_SINGULAR) {
        g_err.clear();
        rc = 0;
        if (inner.empty()) {
          if (!conv32) break;
          conv32 = false;
          continue;
        }
      }
      if (rc) return rc;
      for (double e : inner) hist.push_back(e * rs / bnorm);
      total += (int)inner.size();
      // the Krylov blocks are free between cycles: ws[0] keeps X for an undo
      // (restored by copy, so a correction that is not finite is undone too)
      if (was32) {
        BDDA_HIP(hipMemcpyAsync(Rsave, R64, b64, hipMemcpyDeviceToDevice, s));
        BDDA_HIP(hipMemcpyAsync(ws[0], X, b64, hipMemcpyDeviceToDevice, s));
      }
      launch_axpy_norm<double2>(X, X, make_double2(1.0, 0.0), D, n, nullptr, dot_parts(), s);
      BDDA_HIP(hipGetLastError());
      if ((rc = apply_operator<double2>(p, X, tmp, L, true, s))) return rc;
      double r2new;
      if ((rc = axpy_norm_host<double2>(c, R64, Bj, zc(-1.0), tmp, &r2new))) return rc;
      if (was32 && !(r2new < r2)) {  // the cycle made things worse: undo it
        BDDA_HIP(hipMemcpyAsync(X, ws[0], b64, hipMemcpyDeviceToDevice, s));
        BDDA_HIP(hipMemcpyAsync(R64, Rsave, b64, hipMemcpyDeviceToDevice, s));
        r2new = r2;
      }
      // the next cycle restarts the Krylov space from the true residual: in
      // complex64 again if this one was a complex64 cycle that reached its
      // tolerance (not stopped by its check) and at least halved the
      // residual, else in complex128
      if (!(was32 && conv32 && r2new <= 0.25 * r2)) {
        if (was32) flags |= 2;
        conv32 = false;
      }
      r2 = r2new;
      res = std::sqrt(r2) / bnorm;
      if (!hist.empty()) hist.back() = res;
      if (inner.empty()) break;
    }
    // a guess that already meets the tolerance needs no cycle: its true
    // residual is the solve's one history entry (an empty history reports
    // a solve that never ran, err_fin = inf, as the reference does)
    if (hist.empty() && res < tol) hist.push_back(res);
    it = total > 0 ? total - 1 : 0;
    err = res;
    flags |= std::min(cycles, 63) << 2;
  }
  if (true_res) {
    if (precision == BDDA_PREC_MIXED) {
      *true_res = err;
    } else {
      if ((rc = apply_operator<double2>(p, X, tmp, L, true, s))) return rc;
      double r2 = 0.0;
      if ((rc = axpy_norm_host<double2>(c, tmp, Bj, zc(-1.0), tmp, &r2))) return rc;
      *true_res = std::sqrt(r2) / bnorm;
    }
  }
  BDDA_HIP(hipStreamSynchronize(s));
  if (iter_fin) *iter_fin = it;
  if (err_fin) *err_fin = err;
  if (err_hist)
    for (size_t i = 0; i < hist.size() && (int)i < std::max(itermax, 1); ++i) err_hist[i] = hist[i];
  if (flags_out) *flags_out = flags | ((int)std::min<size_t>(hist.size(), 0x7fffff) << 8);
  return 0;
}

int bdda_true_residual(bdda_plan* p, const void* B_dev, const void* X_dev, int L, double* res, void* stream) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  if (!p->green_ready || !p->target_ready) {
    set_error("bdda_true_residual: call bdda_green_init and bdda_set_target first");
    return BDDA_E_STATE;
  }
  if (L < 1 || !B_dev || !X_dev || !res) {
    set_error("bdda_true_residual: bad arguments");
    return BDDA_E_ARG;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  Ctx c{p, L, 3 * p->N, s, nullptr};
  const size_t b64 = (size_t)c.rows * L * sizeof(double2);
  int rc = p->Bj64.ensure(b64);
  if (!rc) rc = p->blk[8].ensure(b64);
  if (rc) return rc;
  double2* Bj = p->Bj64.as<double2>();
  double2* tmp = p->blk[8].as<double2>();
  launch_scale_rows<double2>(Bj, (const double2*)B_dev, p->jpre64.as<double2>(), c.rows, L, s);
  BDDA_HIP(hipGetLastError());
  double bn2 = 0.0, r2 = 0.0;
  if ((rc = frob2(c, Bj, &bn2))) return rc;
  if ((rc = apply_operator<double2>(p, (const double2*)X_dev, tmp, L, true, s))) return rc;
  if ((rc = axpy_norm_host<double2>(c, tmp, Bj, zc(-1.0), tmp, &r2))) return rc;
  *res = bn2 > 0.0 ? std::sqrt(r2 / bn2) : 0.0;
  return 0;
}

int bdda_solve_batch(int nb, bdda_plan* const* plans, int solver, int precision, const void* const* B,
                     const void* const* X0, void* const* X, const int* L, double tol, int itermax, int refine_every,
                     int nstreams, int* iter_fin, double* err_fin, double* true_res, double* err_hist,
                     int* flags_out) {
  g_err.clear();
  if (nb < 1 || !plans || !B || !X || !L || nstreams < 1) {
    set_error("bdda_solve_batch: need nb >= 1 instances, plans, B, X, L and nstreams >= 1");
    return BDDA_E_ARG;
  }
  std::vector<const bdda_plan*> seen(plans, plans + nb);
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end() || seen.front() == nullptr) {
    set_error("bdda_solve_batch: every instance needs its own (non-null) plan");
    return BDDA_E_ARG;
  }
  // largest instances first: the workers take them in this order
  std::vector<int> order(nb);
  for (int i = 0; i < nb; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (double)plans[a]->N * L[a] > (double)plans[b]->N * L[b];
  });
  const int nthreads = std::min(std::min(nstreams, nb), 64);
  int device = 0;
  BDDA_HIP(hipGetDevice(&device));
  std::vector<hipStream_t> streams(nthreads, nullptr);
  for (int t = 0; t < nthreads; ++t) {
    if (hipStreamCreateWithFlags(&streams[t], hipStreamNonBlocking) != hipSuccess) {
      for (int u = 0; u < t; ++u) (void)hipStreamDestroy(streams[u]);
      set_error("bdda_solve_batch: hipStreamCreateWithFlags failed");
      return BDDA_E_HIP;
    }
  }
  std::atomic<int> next{0};
  std::vector<int> rcs(nb, 0);
  std::vector<std::string> errs(nb);
  const int hstride = std::max(itermax, 1);
  auto worker = [&](int t) {
    (void)hipSetDevice(device);
    for (int k = next++; k < nb; k = next++) {
      const int i = order[k];
      rcs[i] = bdda_solve(plans[i], solver, precision, B[i], X0 ? X0[i] : nullptr, X[i], L[i], tol, itermax,
                          refine_every, iter_fin ? iter_fin + i : nullptr, err_fin ? err_fin + i : nullptr,
                          true_res ? true_res + i : nullptr, err_hist ? err_hist + (size_t)i * hstride : nullptr,
                          flags_out ? flags_out + i : nullptr, streams[t]);
      if (rcs[i]) errs[i] = bdda_last_error();
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker, t);
  worker(0);
  for (auto& th : pool) th.join();
  for (int t = 0; t < nthreads; ++t) (void)hipStreamDestroy(streams[t]);
  for (int i = 0; i < nb; ++i) {
    if (rcs[i]) {
      set_error("bdda_solve_batch: instance " + std::to_string(i) + ": " + errs[i]);
      return rcs[i];
    }
  }
  return 0;
}

int bdda_project_guess(bdda_plan* p, const void* B_dev, void* X0_dev, int L, double* res_before,
                       double* res_after, void* stream) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  if (!p->green_ready || !p->target_ready) {
    set_error("bdda_project_guess: call bdda_green_init and bdda_set_target first");
    return BDDA_E_STATE;
  }
  if (L < 1 || L > kMaxL || !B_dev || !X0_dev) {
    set_error("bdda_project_guess: bad arguments (1 <= L <= 4096, B and X0 on the device)");
    return BDDA_E_ARG;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  Ctx c{p, L, 3 * p->N, s, nullptr};
  const size_t b64 = (size_t)c.rows * L * sizeof(double2);
  int rc = p->Bj64.ensure(b64);
  for (int i = 0; i < 3 && !rc; ++i) rc = p->blk[i].ensure(b64);
  if (rc) return rc;
  double2* X0 = (double2*)X0_dev;
  double2* Bj = p->Bj64.as<double2>();
  double2* W = p->blk[0].as<double2>();
  double2* Y = p->blk[1].as<double2>();
  double2* R = p->blk[2].as<double2>();
  launch_scale_rows<double2>(Bj, (const double2*)B_dev, p->jpre64.as<double2>(), c.rows, L, s);
  BDDA_HIP(hipGetLastError());
  if ((rc = apply_operator<double2>(p, X0, W, L, true, s))) return rc;  // W = jpre (A X0)
  double bn2 = 0.0, r0 = 0.0, r1 = 0.0;
  if ((rc = frob2(c, Bj, &bn2))) return rc;
  if ((rc = axpy_norm_host<double2>(c, R, Bj, zc(-1.0), W, &r0))) return rc;
  // C = argmin ||Bj - W C||_F : (W^H W) C = W^H Bj, with a relative ridge
  // of 1e-13 tr(W^H W) / L against nearly dependent guesses
  std::vector<zc> G, h;
  std::vector<int> piv;
  if ((rc = gram_to_host(c, W, W, true, false, G))) return rc;
  if ((rc = gram_to_host(c, W, Bj, true, false, h))) return rc;
  double tr = 0.0;
  for (int i = 0; i < L; ++i) tr += G[(size_t)i * L + i].real();
  for (int i = 0; i < L; ++i) G[(size_t)i * L + i] += 1e-13 * tr / L;
  if (!(tr > 0.0) || !la::lu_factor(G, piv, L)) return singular("guess projection (W^H W)");
  la::lu_solve(G, piv, L, h, L);
  // X0 <- X0 C ; residual Bj - W C
  if ((rc = update<double2>(c, Y, nullptr, 0.0, X0, nullptr, 0.0, h, 0, nullptr, 0.0))) return rc;
  BDDA_HIP(hipMemcpyAsync(X0, Y, b64, hipMemcpyDeviceToDevice, s));
  std::vector<zc> negC(h);
  for (auto& v : negC) v = -v;
  if ((rc = update<double2>(c, R, Bj, 1.0, W, nullptr, 0.0, negC, 1, nullptr, 0.0))) return rc;
  if ((rc = frob2(c, R, &r1))) return rc;
  BDDA_HIP(hipStreamSynchronize(s));
  const double bn = std::sqrt(bn2);
  if (res_before) *res_before = bn > 0 ? std::sqrt(r0) / bn : 0.0;
  if (res_after) *res_after = bn > 0 ? std::sqrt(r1) / bn : 0.0;
  return 0;
}

int bdda_rhs(bdda_plan* p, const double* u, const void* e0, double k, int L, void* B_dev, void* stream) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  if (!p->target_ready || !p->has_pos || L < 1 || !u || !e0 || !B_dev) {
    set_error("bdda_rhs: bad arguments or no target");
    return BDDA_E_ARG;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  DevBuf du, de;
  int rc = du.ensure(L * 3 * sizeof(double));
  if (!rc) rc = de.ensure(L * 3 * sizeof(double2));
  if (rc) return rc;
  BDDA_HIP(hipMemcpy(du.p, u, L * 3 * sizeof(double), hipMemcpyHostToDevice));
  BDDA_HIP(hipMemcpy(de.p, e0, L * 3 * sizeof(double2), hipMemcpyHostToDevice));
  launch_rhs((double2*)B_dev, p->pos.as<double>(), du.as<double>(), de.as<double2>(), k, p->N, L, s);
  BDDA_HIP(hipGetLastError());
  BDDA_HIP(hipStreamSynchronize(s));
  return 0;
}

int bdda_observables(bdda_plan* p, const void* X_dev, int L, const double* geo, const void* epsfac, double k,
                     void* out_host, void* stream) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  if (!p->target_ready || !p->has_pos || L < 1 || !X_dev || !geo || !epsfac || !out_host) {
    set_error("bdda_observables: bad arguments or no target");
    return BDDA_E_ARG;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int np = obs_parts();
  DevBuf dg, dfac;
  int rc = dg.ensure(L * 24 * sizeof(double));
  if (!rc) rc = dfac.ensure(3 * sizeof(double2));
  if (!rc) rc = p->obs_part.ensure((size_t)np * L * 5 * sizeof(double2));
  if (rc) return rc;
  BDDA_HIP(hipMemcpy(dg.p, geo, L * 24 * sizeof(double), hipMemcpyHostToDevice));
  BDDA_HIP(hipMemcpy(dfac.p, epsfac, 3 * sizeof(double2), hipMemcpyHostToDevice));
  launch_observables((const double2*)X_dev, p->pos.as<double>(), dg.as<double>(), dfac.as<double2>(), k, p->N, L,
                     p->obs_part.as<double2>(), np, s);
  BDDA_HIP(hipGetLastError());
  std::vector<double2> part((size_t)np * L * 5);
  BDDA_HIP(hipMemcpyAsync(part.data(), p->obs_part.p, part.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
  BDDA_HIP(hipStreamSynchronize(s));
  double2* out = reinterpret_cast<double2*>(out_host);
  for (int l = 0; l < L; ++l)
    for (int t = 0; t < 5; ++t) {
      double sx = 0.0, sy = 0.0;
      for (int w = 0; w < np; ++w) {
        sx += part[((size_t)w * L + l) * 5 + t].x;
        sy += part[((size_t)w * L + l) * 5 + t].y;
      }
      out[l * 5 + t] = make_double2(sx, sy);
    }
  return 0;
}

int bdda_observables_noalloc(bdda_plan* p, const void* X_dev, int L, const double* geo_dev, const void* epsfac_dev,
                             double k, void* out_dev, void* stream) {
  if (int rc = check_plan(p)) return rc;
  g_err.clear();
  if (!p->target_ready || !p->has_pos || !X_dev || !geo_dev || !epsfac_dev || !out_dev) {
    set_error("bdda_observables_noalloc: bad arguments or no target");
    return BDDA_E_ARG;
  }
  if (L < 1 || L > kNoAllocMaxL) {
    set_error("bdda_observables_noalloc: L = " + std::to_string(L) + " is outside 1.." +
              std::to_string(kNoAllocMaxL) + " (the size of its fixed scratch buffer)");
    return BDDA_E_ARG;
  }
  const int np = obs_parts();
  if ((size_t)np * L * 5 > kObsScratchElems) {
    set_error("bdda_observables_noalloc: partial sums exceed the scratch buffer");
    return BDDA_E_ARG;
  }
  void* part = nullptr;
  if (int rc = scratch_ptr(&part)) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // every partial slot (workgroup < np, l < L) is written by the kernel
  // before the reduction reads it, so the scratch needs no clearing
  launch_observables((const double2*)X_dev, p->pos.as<double>(), geo_dev, (const double2*)epsfac_dev, k, p->N, L,
                     (double2*)part, np, s);
  BDDA_HIP(hipGetLastError());
  launch_obs_reduce((const double2*)part, np, L, (double2*)out_dev, s);
  BDDA_HIP(hipGetLastError());
  return 0;
}

int bdda_nearest_distance(const double* grid, int64_t n, const double* surf, int64_t P, double* dist_host) {
  g_err.clear();
  if (!grid || !surf || !dist_host || n < 1 || P < 1) {
    set_error("bdda_nearest_distance: bad arguments");
    return BDDA_E_ARG;
  }
  DevBuf dgrid, dsurf, ddist;
  int rc = dgrid.ensure(n * 3 * sizeof(double));
  if (!rc) rc = dsurf.ensure(P * 3 * sizeof(double));
  if (!rc) rc = ddist.ensure(n * sizeof(double));
  if (rc) return rc;
  BDDA_HIP(hipMemcpy(dgrid.p, grid, n * 3 * sizeof(double), hipMemcpyHostToDevice));
  BDDA_HIP(hipMemcpy(dsurf.p, surf, P * 3 * sizeof(double), hipMemcpyHostToDevice));
  launch_voxel_dist(dgrid.as<double>(), n, dsurf.as<double>(), P, ddist.as<double>(), 0);
  BDDA_HIP(hipGetLastError());
  BDDA_HIP(hipMemcpy(dist_host, ddist.p, n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int bdda_inside_mask(const double* dist, int nx, int ny, int nz, double thresh, uint8_t* inside_host) {
  g_err.clear();
  if (!dist || !inside_host || nx < 1 || ny < 1 || nz < 1) {
    set_error("bdda_inside_mask: bad arguments");
    return BDDA_E_ARG;
  }
  const long long n = (long long)nx * ny * nz;
  DevBuf ddist, dins;
  int rc = ddist.ensure(n * sizeof(double));
  if (!rc) rc = dins.ensure(n);
  if (rc) return rc;
  BDDA_HIP(hipMemcpy(ddist.p, dist, n * sizeof(double), hipMemcpyHostToDevice));
  launch_voxel_inside(ddist.as<double>(), thresh, nx, ny, nz, dins.as<unsigned char>(), 0);
  BDDA_HIP(hipGetLastError());
  BDDA_HIP(hipMemcpy(inside_host, dins.p, n, hipMemcpyDeviceToHost));
  return 0;
}

// ----------------------------------------------------- learned warm start
int bdda_gemm_bf16(int batch, int M, int N, int K, float alpha, const float* A, int64_t sAb, int64_t sAi,
                   int64_t sAk, const float* B, int64_t sBb, int64_t sBk, int64_t sBj, float beta, float* C,
                   int64_t sCb, int64_t sCi, int64_t sCj, void* stream) {
  return bdda_gemm_bf16_ex(batch, M, N, K, alpha, A, sAb, sAi, sAk, B, sBb, sBk, sBj, beta, C, sCb, sCi, sCj,
                           nullptr, 0, stream);
}

int bdda_gemm_bf16_ex(int batch, int M, int N, int K, float alpha, const float* A, int64_t sAb, int64_t sAi,
                      int64_t sAk, const float* B, int64_t sBb, int64_t sBk, int64_t sBj, float beta, float* C,
                      int64_t sCb, int64_t sCi, int64_t sCj, const float* bias, int act, void* stream) {
  g_err.clear();
  if (batch < 0 || M < 0 || N < 0 || K < 0 || !A || !B || !C || (act != 0 && act != 1)) {
    set_error("bdda_gemm_bf16: bad arguments");
    return BDDA_E_ARG;
  }
  if (batch == 0 || M == 0 || N == 0) return 0;
  GemmArgs g{0, M, N, K, alpha, beta, A, sAb, sAi, sAk, B, sBb, sBk, sBj, C, sCb, sCi, sCj, bias, act};
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    g.batch = std::min(65535, batch - b0);
    g.A = A + (long long)b0 * sAb;
    g.B = B + (long long)b0 * sBb;
    g.C = C + (long long)b0 * sCb;
    launch_gemm_bf16(g, s);
    BDDA_HIP(hipGetLastError());
  }
  return 0;
}

int bdda_fno_inputs(float* out, const uint8_t* occ, const int* sample_inst, const float* desc, int G, int64_t nsamp,
                    void* stream) {
  g_err.clear();
  if (!out || !occ || !sample_inst || !desc || G < 1 || nsamp < 0) {
    set_error("bdda_fno_inputs: bad arguments");
    return BDDA_E_ARG;
  }
  launch_fno_inputs(out, occ, sample_inst, desc, G, nsamp, reinterpret_cast<hipStream_t>(stream));
  BDDA_HIP(hipGetLastError());
  return 0;
}

int bdda_cube_exchange(float* cube, void* X, const int64_t* cells, const void* alpha, int64_t N, int L, int G,
                       int to_cube, void* stream) {
  g_err.clear();
  if (!cube || !X || !cells || !alpha || N < 0 || L < 1 || G < 1) {
    set_error("bdda_cube_exchange: bad arguments");
    return BDDA_E_ARG;
  }
  launch_cube_exchange(cube, (double2*)X, (const long long*)cells, (const double2*)alpha, N, L, G, to_cube,
                       reinterpret_cast<hipStream_t>(stream));
  BDDA_HIP(hipGetLastError());
  return 0;
}

int bdda_masked_l2(const float* pred, const float* target, const uint8_t* occ, const int* sample_inst, int C,
                   int64_t S, float* loss, float* grad, int nsamp, void* stream) {
  g_err.clear();
  if (!pred || !target || !occ || !sample_inst || !loss || nsamp < 1) {
    set_error("bdda_masked_l2: bad arguments");
    return BDDA_E_ARG;
  }
  launch_masked_l2(pred, target, occ, sample_inst, C, S, loss, grad, nsamp, reinterpret_cast<hipStream_t>(stream));
  BDDA_HIP(hipGetLastError());
  return 0;
}

int bdda_gelu(const float* x, float* y, int64_t n, void* stream) {
  launch_gelu(x, y, n, reinterpret_cast<hipStream_t>(stream));
  BDDA_HIP(hipGetLastError());
  return 0;
}

int bdda_gelu_bwd(const float* x, const float* dy, float* dx, int64_t n, void* stream) {
  launch_gelu_bwd(x, dy, dx, n, reinterpret_cast<hipStream_t>(stream));
  BDDA_HIP(hipGetLastError());
  return 0;
}

int bdda_bias_add(float* y, const float* b, int64_t B, int C, int64_t S, void* stream) {
  launch_bias_add(y, b, B, C, S, reinterpret_cast<hipStream_t>(stream));
  BDDA_HIP(hipGetLastError());
  return 0;
}

int bdda_fno_ycombine(const float* R, float* xr, float* xi, int64_t nq, int K, void* stream) {
  g_err.clear();
  if (!R || !xr || !xi || nq < 0 || K < 1) {
    set_error("bdda_fno_ycombine: bad arguments");
    return BDDA_E_ARG;
  }
  if (nq == 0) return 0;
  launch_fno_ycombine(R, xr, xi, nq, K, reinterpret_cast<hipStream_t>(stream));
  BDDA_HIP(hipGetLastError());
  return 0;
}

int bdda_bias_grad(const float* dy, float* db, int64_t B, int C, int64_t S, void* stream) {
  launch_bias_grad(dy, db, B, C, S, reinterpret_cast<hipStream_t>(stream));
  BDDA_HIP(hipGetLastError());
  return 0;
}

int bdda_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps,
              int step, void* stream) {
  if (step < 1) {
    set_error("bdda_adam: step counts from 1");
    return BDDA_E_ARG;
  }
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  launch_adam(p, g, m, v, n, lr, b1, b2, eps, c1, c2, reinterpret_cast<hipStream_t>(stream));
  BDDA_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
