// Batched line FFTs for the Goodman doubled-grid convolution on gfx950:
// kernel selection, the runtime-radix launches, the Green tensor and the
// radix plans.  The kernels themselves are in fft_impl.h; the register
// kernels are instantiated in fft_len_*.hip.
#include "fft_impl.h"

namespace bdda {

std::atomic<bool> g_force_generic{false};
std::atomic<bool> g_xconv_off{false};
std::atomic<int> g_xconv_split{0};

BDDA_REG2_LENGTHS(BDDA_REG2_EXTERN)

// Lengths with a two-step register kernel (m = R1 * R2, BDDA_REG2_LENGTHS);
// others use the runtime-radix Stockham kernel.  A register kernel needs
// Q * max(R1, R2) threads for one point, which must stay within its
// 512-thread bound; otherwise the length also goes to the Stockham kernel.
template <typename V>
static bool try_reg2(const LinePass& lp, hipStream_t s) {
  switch (lp.m) {
#define BDDA_REG2_CASE(m, r1, r2)                      \
  case m:                                              \
    if (lp.Q * (r1 > r2 ? r1 : r2) > 512) return false; \
    launch_reg2<V, r1, r2>(lp, s);                     \
    return true;
    BDDA_REG2_LENGTHS(BDDA_REG2_CASE)
#undef BDDA_REG2_CASE
    default:
      return false;
  }
}

template <typename V>
void launch_line_pass(const LinePass& lp0, hipStream_t s) {
  LinePass lp = lp0;
  if (lp.m * lp.Q > kTile) {
    set_error("line FFT tile too large (m*Q must be <= 4096)");
    return;
  }
  if (!g_force_generic && try_reg2<V>(lp, s)) return;
  const int TPC = (lp.m + kVPT - 1) / kVPT;
  if (lp.Q * TPC > 512) {
    set_error("line FFT tile too large (m*Q must be <= 4096)");
    return;
  }
  int P = std::min(kTile / (lp.m * lp.Q), 512 / (lp.Q * TPC));
  if (P < 1) P = 1;
  if (P > lp.npts) P = lp.npts;
  lp.P = P;
  const int nthr = P * lp.Q * TPC;
  if (nthr > 512) {
    set_error("line FFT tile too large (launch exceeds 512 threads)");
    return;
  }
  const size_t lds = ((size_t)lp.m * P * lp.Q + lp.m) * sizeof(V);
  dim3 grid((lp.npts + P - 1) / P, lp.nlo > 0 ? lp.nlo : 1);
  switch (lp.mode) {
    case PASS_PLAIN: launch_mode<V, PASS_PLAIN>(lp, grid, nthr, lds, s); break;
    case PASS_SCATTER_Z: launch_mode<V, PASS_SCATTER_Z>(lp, grid, nthr, lds, s); break;
    case PASS_CONV_X: launch_mode<V, PASS_CONV_X>(lp, grid, nthr, lds, s); break;
    case PASS_GATHER_Z: launch_mode<V, PASS_GATHER_Z>(lp, grid, nthr, lds, s); break;
    case PASS_CUBOID_Z: launch_mode<V, PASS_CUBOID_Z>(lp, grid, nthr, lds, s); break;
    case PASS_SCATTER_Z64:
    case PASS_GATHER_Z64:
      if constexpr (std::is_same<V, float2>::value) {
        if (lp.mode == PASS_SCATTER_Z64)
          launch_mode<V, PASS_SCATTER_Z64>(lp, grid, nthr, lds, s);
        else
          launch_mode<V, PASS_GATHER_Z64>(lp, grid, nthr, lds, s);
      } else {
        set_error("line FFT: complex128 boundary modes need a complex64 transform");
      }
      break;
    default: launch_mode<V, PASS_CROP_Z>(lp, grid, nthr, lds, s); break;
  }
}

void set_fft_path(int path) {
  g_force_generic = path == 1;
  g_xconv_off = path == 2;
  g_xconv_split = path == 3 ? 1 : (path == 4 ? 2 : 0);
}

template void launch_line_pass<double2>(const LinePass&, hipStream_t);
template void launch_line_pass<float2>(const LinePass&, hipStream_t);

// ------------------------------------------------------------ Green tensor
__device__ __forceinline__ int green_disp(int I, int N, int M, bool& gap) {
  if (I <= N - 1) return I;
  if (I >= M - N + 1) return I - M;
  gap = true;
  return 0;
}

// A'(r) = e^{ikr}/r [k^2 (rr - I) + (1 - ikr)/r^2 (I - 3 rr)] on the m-grid,
// 6 unique components (xx, xy, xz, yy, yz, zz); origin and gap cells zero.
__global__ void k_green_build(double2* out, int nx, int ny, int nz, int mx, int my, int mz, double lf,
                              double k) {
  const long long total = (long long)mx * my * mz;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    const int iz = id % mz;
    const int iy = (id / mz) % my;
    const int ix = id / ((long long)mz * my);
    bool gap = false;
    const double dx = green_disp(ix, nx, mx, gap) * lf;
    const double dy = green_disp(iy, ny, my, gap) * lf;
    const double dz = green_disp(iz, nz, mz, gap) * lf;
    double2* o = out + id * 6;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (gap || r == 0.0) {
      for (int c = 0; c < 6; ++c) o[c] = make_double2(0.0, 0.0);
      continue;
    }
    const double u[3] = {dx / r, dy / r, dz / r};
    double sn, cs;
    sincos(k * r, &sn, &cs);
    const double2 ph = make_double2(cs / r, sn / r);                 // e^{ikr}/r
    const double2 c2 = make_double2(1.0 / (r * r), -k * r / (r * r));  // (1 - ikr)/r^2
    const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
    for (int c = 0; c < 6; ++c) {
      const double rr = u[ia[c]] * u[ib[c]];
      const double dl = ia[c] == ib[c] ? 1.0 : 0.0;
      const double t1 = k * k * (rr - dl);
      const double t2 = dl - 3.0 * rr;
      const double2 term = make_double2(t1 + c2.x * t2, c2.y * t2);
      o[c] = make_double2(ph.x * term.x - ph.y * term.y, ph.x * term.y + ph.y * term.x);
    }
  }
}

template <typename V>
__global__ void k_green_halve(const double2* full, V* half, int mx, int my, int mz, int hx, int hy, int hz) {
  const long long total = (long long)hx * hy * hz * 6;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    // half spectrum stored x-innermost: [(qy*hz + qz)*hx + qx][6]
    const int c = id % 6;
    const long long t = id / 6;
    const int qx = t % hx, qz = (t / hx) % hz, qy = t / ((long long)hx * hz);
    const double2 g = full[(((long long)qx * my + qy) * mz + qz) * 6 + c];
    V v;
    v.x = g.x;
    v.y = g.y;
    half[id] = v;
  }
}

__global__ void k_green_expand(const double2* half, double2* full, int mx, int my, int mz, int hx, int hy,
                               int hz) {
  const long long total = (long long)mx * my * mz;
  const int comp[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    const int qz = id % mz, qy = (id / mz) % my, qx = id / ((long long)mz * my);
    int ax = qx, ay = qy, az = qz;
    double s[3] = {1, 1, 1};
    if (ax > mx - ax) { ax = mx - ax; s[0] = -1; }
    if (ay > my - ay) { ay = my - ay; s[1] = -1; }
    if (az > mz - az) { az = mz - az; s[2] = -1; }
    const double2* g = half + (((long long)ay * hz + az) * hx + ax) * 6;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        const double sg = a == b ? 1.0 : s[a] * s[b];
        const double2 v = g[comp[a * 3 + b]];
        full[id * 9 + a * 3 + b] = make_double2(v.x * sg, v.y * sg);
      }
  }
}

__global__ void k_green_sym9(const double2* g6, double2* g9, long long M) {
  const int comp[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < M * 9;
       id += (long long)gridDim.x * blockDim.x)
    g9[id] = g6[(id / 9) * 6 + comp[id % 9]];
}

void launch_green_sym9(const double2* g6, double2* g9, long long M, hipStream_t s) {
  hipLaunchKernelGGL(k_green_sym9, dim3(2048), dim3(256), 0, s, g6, g9, M);
}

template <typename V>
void launch_green_build(V* out, const AxisPlan* ax, double lf, double k, hipStream_t s) {
  hipLaunchKernelGGL(k_green_build, dim3(2048), dim3(256), 0, s, reinterpret_cast<double2*>(out), ax[0].n,
                     ax[1].n, ax[2].n, ax[0].m, ax[1].m, ax[2].m, lf, k);
}
template void launch_green_build<double2>(double2*, const AxisPlan*, double, double, hipStream_t);

template <typename V>
void launch_green_halve(const double2* full, V* half, const AxisPlan* ax, hipStream_t s) {
  hipLaunchKernelGGL((k_green_halve<V>), dim3(2048), dim3(256), 0, s, full, half, ax[0].m, ax[1].m, ax[2].m,
                     ax[0].h, ax[1].h, ax[2].h);
}
template void launch_green_halve<double2>(const double2*, double2*, const AxisPlan*, hipStream_t);
template void launch_green_halve<float2>(const double2*, float2*, const AxisPlan*, hipStream_t);

void launch_green_expand(const double2* half, double2* full, const AxisPlan* ax, const int*, hipStream_t s) {
  hipLaunchKernelGGL(k_green_expand, dim3(2048), dim3(256), 0, s, half, full, ax[0].m, ax[1].m, ax[2].m,
                     ax[0].h, ax[1].h, ax[2].h);
}

// ------------------------------------------------------------ radix plans
int smallest_5smooth_at_least(int v) {
  for (int m = v < 1 ? 1 : v;; ++m) {
    int t = m;
    for (int p : {2, 3, 5})
      while (t % p == 0) t /= p;
    if (t == 1) return m;
  }
}

// Register-kernel lengths and their first radix.
static const int kRegLengths[][2] = {
#define BDDA_REG_ENTRY(m, r1, r2) {m, r1},
    BDDA_REG2_LENGTHS(BDDA_REG_ENTRY)
#undef BDDA_REG_ENTRY
};

int conv_length(int n, int axis) {
  const int two = 2 * n;
  if (smallest_5smooth_at_least(two) == two) return two;
  const int lo = 2 * n - 1;
  const int m = smallest_5smooth_at_least(lo);
  if (axis != 0) return m;
  // x: the dedicated x-pass kernel (k_xconv) needs a register length with an even
  // first radix; take the smallest such length >= 2N - 1 when there is one
  // (45, 75, 81, 135 and 225 become 48, 80, 90, 144 and 240)
  int best = 0;
  for (const auto& e : kRegLengths)
    if (e[0] >= lo && e[1] % 2 == 0 && (best == 0 || e[0] < best)) best = e[0];
  return best ? best : m;
}

bool factor_length(int m, AxisPlan& ax) {
  ax.m = m;
  ax.h = m / 2 + 1;
  ax.nrad = 0;
  int t = m;
  auto take = [&](int r) {
    while (t % r == 0) {
      if (ax.nrad >= kMaxRadix) return false;
      ax.rad[ax.nrad++] = r;
      t /= r;
    }
    return true;
  };
  if (!take(8) || !take(4) || !take(2) || !take(3) || !take(5)) return false;
  return t == 1;
}

}  // namespace bdda
