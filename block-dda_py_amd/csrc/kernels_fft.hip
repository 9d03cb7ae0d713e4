// Batched line FFTs for the Goodman doubled-grid convolution on gfx950.
//
// Every launch transforms a tile of P points x Q columns x m elements held in
// LDS ([m][P*Q], columns innermost so that 64 consecutive lanes touch 64
// consecutive LDS words).  The transform is a mixed-radix (8, 4, 2, 3, 5)
// Stockham autosort FFT done in place in LDS: each pass stages its butterflies
// in registers, synchronises, and writes the permuted outputs back.
//
// Global layouts (complex units, columns innermost):
//   occupied block  (3N, L) row-major, row = 3m + c
//   T1  [lo][ix][iy][kz][c][li]   (Nx, Ny, Mz, 3, Lb)  after the z pass
//   T2  [lo][ix][ky][kz][c][li]   (Nx, My, Mz, 3, Lb)  after the y pass
// The x pass (forward x, 3x3 Green multiply, inverse x) runs in place on T2.
#include "internal.h"

namespace bdda {

template <typename V> struct RealOf;
template <> struct RealOf<double2> { using T = double; };
template <> struct RealOf<float2> { using T = float; };

template <typename V>
__device__ __forceinline__ V cx(typename RealOf<V>::T x, typename RealOf<V>::T y) {
  V v;
  v.x = x;
  v.y = y;
  return v;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return cx<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return cx<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  return cx<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V> __device__ __forceinline__ V cscale(V a, typename RealOf<V>::T s) {
  return cx<V>(a.x * s, a.y * s);
}
// a * (-i) for the forward transform, a * (+i) for the inverse one
template <bool INV, typename V> __device__ __forceinline__ V rotq(V a) {
  return INV ? cx<V>(-a.y, a.x) : cx<V>(a.y, -a.x);
}

// ------------------------------------------------------------- butterflies
template <int R, bool INV, typename V> struct Dft;

template <bool INV, typename V> struct Dft<2, INV, V> {
  __device__ __forceinline__ static void run(V* a) {
    V t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <bool INV, typename V> struct Dft<4, INV, V> {
  __device__ __forceinline__ static void run(V* a) {
    V t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    V t2 = cadd(a[1], a[3]), t3 = rotq<INV>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <bool INV, typename V> struct Dft<8, INV, V> {
  __device__ __forceinline__ static void run(V* a) {
    using T = typename RealOf<V>::T;
    const T h = (T)0.70710678118654752440;
    V e[4] = {a[0], a[2], a[4], a[6]};
    V o[4] = {a[1], a[3], a[5], a[7]};
    Dft<4, INV, V>::run(e);
    Dft<4, INV, V>::run(o);
    // o1 *= W8^1, o2 *= W8^2, o3 *= W8^3   (W8 = exp(-+ i pi/4))
    V o1 = o[1], o3 = o[3];
    if (!INV) {
      o[1] = cx<V>((o1.x + o1.y) * h, (o1.y - o1.x) * h);
      o[3] = cx<V>((o3.y - o3.x) * h, -(o3.x + o3.y) * h);
    } else {
      o[1] = cx<V>((o1.x - o1.y) * h, (o1.x + o1.y) * h);
      o[3] = cx<V>(-(o3.x + o3.y) * h, (o3.x - o3.y) * h);
    }
    o[2] = rotq<INV>(o[2]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = cadd(e[k], o[k]);
      a[k + 4] = csub(e[k], o[k]);
    }
  }
};

template <bool INV, typename V> struct Dft<3, INV, V> {
  __device__ __forceinline__ static void run(V* a) {
    using T = typename RealOf<V>::T;
    const T s = (T)0.86602540378443864676;
    V t = cadd(a[1], a[2]);
    V d = cscale(rotq<INV>(csub(a[1], a[2])), s);
    V m = cx<V>(a[0].x - (T)0.5 * t.x, a[0].y - (T)0.5 * t.y);
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, d);
    a[2] = csub(m, d);
  }
};

template <bool INV, typename V> struct Dft<5, INV, V> {
  __device__ __forceinline__ static void run(V* a) {
    using T = typename RealOf<V>::T;
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    V t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
    V d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
    V m1 = cx<V>(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
    V m2 = cx<V>(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
    V n1 = rotq<INV>(cx<V>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
    V n2 = rotq<INV>(cx<V>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
    a[0] = cadd(a[0], cadd(t1, t2));
    a[1] = cadd(m1, n1);
    a[4] = csub(m1, n1);
    a[2] = cadd(m2, n2);
    a[3] = csub(m2, n2);
  }
};

// One Stockham pass of radix R over QQ independent columns of length M.
// Requires M*QQ <= VPT*nthr so that every thread stages at most VPT values.
template <int R, bool INV, int VPT, typename V>
__device__ __forceinline__ void stockham_pass(V* buf, int M, int QQ, int Ns, const V* __restrict__ tw,
                                              int tid, int nthr) {
  constexpr int NB = (VPT + R - 1) / R;
  const int nb = M / R;
  const int total = nb * QQ;
  const int ts = M / (Ns * R);
  V v[NB][R];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = tid + i * nthr;
    if (b < total) {
      const int q = b % QQ, j = b / QQ;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[(j + r * nb) * QQ + q];
      const int jm = j % Ns;
      if (jm != 0) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          V w = tw[jm * r * ts];
          if (INV) w.y = -w.y;
          v[i][r] = cmul(v[i][r], w);
        }
      }
      Dft<R, INV, V>::run(v[i]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = tid + i * nthr;
    if (b < total) {
      const int q = b % QQ, j = b / QQ;
      const int jm = j % Ns;
      const int base = (j / Ns) * Ns * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(base + r * Ns) * QQ + q] = v[i][r];
    }
  }
  __syncthreads();
}

template <bool INV, int VPT, typename V>
__device__ void lds_fft(V* buf, const LinePass& lp, int QQ, const V* __restrict__ tw, int tid, int nthr) {
  int Ns = 1;
  for (int p = 0; p < lp.nrad; ++p) {
    const int R = lp.rad[p];
    switch (R) {
      case 8: stockham_pass<8, INV, VPT>(buf, lp.m, QQ, Ns, tw, tid, nthr); break;
      case 4: stockham_pass<4, INV, VPT>(buf, lp.m, QQ, Ns, tw, tid, nthr); break;
      case 2: stockham_pass<2, INV, VPT>(buf, lp.m, QQ, Ns, tw, tid, nthr); break;
      case 3: stockham_pass<3, INV, VPT>(buf, lp.m, QQ, Ns, tw, tid, nthr); break;
      default: stockham_pass<5, INV, VPT>(buf, lp.m, QQ, Ns, tw, tid, nthr); break;
    }
    Ns *= R;
  }
}

// 3x3 symmetric Green multiply at every (qx, point) of an x-line tile.
template <typename V>
__device__ void green_multiply(V* buf, const LinePass& lp, int pt0, int np_here, int QQ, int tid, int nthr) {
  using T = typename RealOf<V>::T;
  const V* __restrict__ G = reinterpret_cast<const V*>(lp.green);
  const int Lb = lp.Lb, Q = lp.Q;
  const int work = lp.m * np_here;
  for (int w = tid; w < work; w += nthr) {
    const int p = w % np_here, qx = w / np_here;
    const int pt = pt0 + p;
    const int ky = pt / lp.mz, kz = pt % lp.mz;
    int ax = qx, ay = ky, az = kz;
    T sx = 1, sy = 1, sz = 1;
    if (ax > lp.mx - ax) { ax = lp.mx - ax; sx = -1; }
    if (ay > lp.my - ay) { ay = lp.my - ay; sy = -1; }
    if (az > lp.mz - az) { az = lp.mz - az; sz = -1; }
    const V* g = G + ((long long)(ax * lp.hy + ay) * lp.hz + az) * 6;
    const V gxx = g[0], gyy = g[3], gzz = g[5];
    const V gxy = cscale(g[1], sx * sy), gxz = cscale(g[2], sx * sz), gyz = cscale(g[4], sy * sz);
    V* row = buf + qx * QQ + p * Q;
    for (int li = 0; li < Lb; ++li) {
      const V x0 = row[li], x1 = row[Lb + li], x2 = row[2 * Lb + li];
      row[li] = cadd(cadd(cmul(gxx, x0), cmul(gxy, x1)), cmul(gxz, x2));
      row[Lb + li] = cadd(cadd(cmul(gxy, x0), cmul(gyy, x1)), cmul(gyz, x2));
      row[2 * Lb + li] = cadd(cadd(cmul(gxz, x0), cmul(gyz, x1)), cmul(gzz, x2));
    }
  }
  __syncthreads();
}

template <typename V, int VPT, int MODE>
__global__ void __launch_bounds__(512) k_line_pass(LinePass lp) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* buf = reinterpret_cast<V*>(smem_raw);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int P = lp.P, Q = lp.Q, QQ = P * Q, M = lp.m;
  const int pt0 = blockIdx.x * P;
  const int lo = blockIdx.y;
  const int np_here = min(P, lp.npts - pt0);
  const int tot = M * QQ;
  const V zero = cx<V>(0, 0);
  const V* __restrict__ src = reinterpret_cast<const V*>(lp.src) + (long long)lo * lp.lo_src;
  const V* __restrict__ tw = reinterpret_cast<const V*>(lp.tw);

  // ---- load (zero-padded to m) -------------------------------------------
  for (int idx = tid; idx < tot; idx += nthr) {
    const int q = idx % Q;
    const int t = idx / Q;
    int p, e;
    if (lp.order_pe) { p = t % P; e = t / P; } else { e = t % M; p = t / M; }
    V v = zero;
    if (e < lp.nin && p < np_here) {
      const int pt = pt0 + p;
      if (MODE == PASS_SCATTER_Z || MODE == PASS_CUBOID_Z) {
        const long long cell = (long long)pt * lp.nz + e;
        const long long m = (MODE == PASS_SCATTER_Z) ? (long long)lp.inv[cell] : cell;
        const int c = q / lp.Lb, l = lo * lp.Lb + q % lp.Lb;
        if (m >= 0 && l < lp.L) v = reinterpret_cast<const V*>(lp.S)[(3 * m + c) * lp.L + l];
      } else {
        const int a = pt / lp.ptb, b = pt % lp.ptb;
        v = src[a * lp.sa + b * lp.sb + e * lp.se + q];
      }
    }
    buf[e * QQ + p * Q + q] = v;
  }
  __syncthreads();

  // ---- transform ------------------------------------------------------------
  if (MODE == PASS_GATHER_Z || MODE == PASS_CROP_Z || (MODE == PASS_PLAIN && lp.inverse)) {
    lds_fft<true, VPT>(buf, lp, QQ, tw, tid, nthr);
  } else {
    lds_fft<false, VPT>(buf, lp, QQ, tw, tid, nthr);
    if (MODE == PASS_CONV_X) {
      green_multiply(buf, lp, pt0, np_here, QQ, tid, nthr);
      lds_fft<true, VPT>(buf, lp, QQ, tw, tid, nthr);
    }
  }

  // ---- store (cropped to nout) ----------------------------------------------
  if (MODE == PASS_GATHER_Z || MODE == PASS_CROP_Z) {
    using T = typename RealOf<V>::T;
    const T sc = (T)lp.scale;
    V* __restrict__ out = reinterpret_cast<V*>(lp.AS);
    const V* __restrict__ S = reinterpret_cast<const V*>(lp.S);
    const V* __restrict__ dg = reinterpret_cast<const V*>(lp.diag);
    const V* __restrict__ jp = reinterpret_cast<const V*>(lp.jpre);
    for (int idx = tid; idx < tot; idx += nthr) {
      const int q = idx % Q;
      const int t = idx / Q;
      const int e = t % M, p = t / M;
      if (e >= lp.nout || p >= np_here) continue;
      const int c = q / lp.Lb, l = lo * lp.Lb + q % lp.Lb;
      if (l >= lp.L) continue;
      const long long cell = (long long)(pt0 + p) * lp.nz + e;
      const V y = cscale(buf[e * QQ + p * Q + q], sc);
      if (MODE == PASS_GATHER_Z) {
        const long long m = lp.inv[cell];
        if (m < 0) continue;
        const long long row = 3 * m + c;
        const V d = cmul(dg[row], S[row * lp.L + l]);
        out[row * lp.L + l] = cmul(jp[row], cadd(d, y));
      } else {
        out[(3 * cell + c) * lp.L + l] = y;
      }
    }
    return;
  }
  V* __restrict__ dst = reinterpret_cast<V*>(lp.dst) + (long long)lo * lp.lo_dst;
  for (int idx = tid; idx < tot; idx += nthr) {
    const int q = idx % Q;
    const int t = idx / Q;
    int p, e;
    if (lp.order_pe) { p = t % P; e = t / P; } else { e = t % M; p = t / M; }
    if (e >= lp.nout || p >= np_here) continue;
    const int pt = pt0 + p;
    const int a = pt / lp.ptb, b = pt % lp.ptb;
    dst[a * lp.da + b * lp.db + e * lp.de + q] = buf[e * QQ + p * Q + q];
  }
}

constexpr int kVPT = 8;
constexpr int kTile = 4096;  // complex values per workgroup tile

template <typename V, int MODE>
static void launch_mode(const LinePass& lp, dim3 grid, int nthr, size_t lds, hipStream_t s) {
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_line_pass<V, kVPT, MODE>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
  hipLaunchKernelGGL((k_line_pass<V, kVPT, MODE>), grid, dim3(nthr), lds, s, lp);
}

template <typename V>
void launch_line_pass(const LinePass& lp0, hipStream_t s) {
  LinePass lp = lp0;
  const int MQ = lp.m * lp.Q;
  int P = kTile / MQ;
  if (P < 1) P = 1;
  if (P > lp.npts) P = lp.npts;
  lp.P = P;
  const int tot = MQ * P;
  int nthr = ((tot + kVPT - 1) / kVPT + 63) / 64 * 64;
  if (nthr < 64) nthr = 64;
  if (nthr > 512 || tot > kVPT * nthr) {
    set_error("line FFT tile too large (m*Q must be <= 4096)");
    return;
  }
  const size_t lds = (size_t)tot * sizeof(V);
  dim3 grid((lp.npts + P - 1) / P, lp.nlo > 0 ? lp.nlo : 1);
  switch (lp.mode) {
    case PASS_PLAIN: launch_mode<V, PASS_PLAIN>(lp, grid, nthr, lds, s); break;
    case PASS_SCATTER_Z: launch_mode<V, PASS_SCATTER_Z>(lp, grid, nthr, lds, s); break;
    case PASS_CONV_X: launch_mode<V, PASS_CONV_X>(lp, grid, nthr, lds, s); break;
    case PASS_GATHER_Z: launch_mode<V, PASS_GATHER_Z>(lp, grid, nthr, lds, s); break;
    case PASS_CUBOID_Z: launch_mode<V, PASS_CUBOID_Z>(lp, grid, nthr, lds, s); break;
    default: launch_mode<V, PASS_CROP_Z>(lp, grid, nthr, lds, s); break;
  }
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
    const int c = id % 6;
    const long long t = id / 6;
    const int qz = t % hz, qy = (t / hz) % hy, qx = t / ((long long)hz * hy);
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
    const double2* g = half + (((long long)ax * hy + ay) * hz + az) * 6;
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
