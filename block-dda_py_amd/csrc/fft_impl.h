#pragma once
// Line-FFT kernels and their launchers, shared by the translation units
// that instantiate them: kernels_fft.hip (dispatch, runtime-radix kernels,
// Green tensor) and fft_len_*.hip (the register kernels, a quarter of the
// lengths each, so that they compile in parallel).  Every kernel instance is
// instantiated in exactly one translation unit.
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
#include <algorithm>
#include <atomic>
#include <mutex>
#include <type_traits>

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
template <typename D, typename S> __device__ __forceinline__ D cvt(S v) {
  D d;
  d.x = v.x;
  d.y = v.y;
  return d;
}
// Type of the occupied block at the scatter / gather boundary: the *_Z64
// modes read and write complex128 around a complex64 transform.
template <typename V, int MODE> struct BoundaryOf { using T = V; };
template <typename V> struct BoundaryOf<V, PASS_SCATTER_Z64> { using T = double2; };
template <typename V> struct BoundaryOf<V, PASS_GATHER_Z64> { using T = double2; };
template <int MODE> constexpr bool is_scatter() { return MODE == PASS_SCATTER_Z || MODE == PASS_SCATTER_Z64; }
template <int MODE> constexpr bool is_gather() { return MODE == PASS_GATHER_Z || MODE == PASS_GATHER_Z64; }
template <int MODE>
constexpr bool is_z_pass() {
  return is_scatter<MODE>() || is_gather<MODE>() || MODE == PASS_CUBOID_Z || MODE == PASS_CROP_Z;
}

// Guarded load returning a value: `c ? p[i] : zero` with two lvalues would
// make the compiler load through a pointer select (a flat load that may hit a
// private copy of zero), so the zero is produced as a value instead.
template <typename V> __device__ __forceinline__ V load_or_zero(bool c, const V* p) {
  V v = cx<V>(0, 0);
  if (c) v = *p;
  return v;
}
// First cell (iz = 0) of z line zl = ix Ny + iy (cells are numbered
// zl Nz + iz).
__device__ __forceinline__ long long zline_cell(const LinePass& lp, int zl) { return (long long)zl * lp.nz; }

// Z line run by point pt of a scatter / gather z pass (an occupied-line list
// when given; the cuboid passes run every line).
template <int MODE>
__device__ __forceinline__ int zline_of(const LinePass& lp, int pt, bool active) {
  if constexpr (is_scatter<MODE>() || is_gather<MODE>())
    return (lp.zlines != nullptr && active) ? lp.zlines[pt] : pt;
  return pt;
}
// Whether the y inverse pass stores element e (y row) of point (ix, kz): T1
// keeps zeros on the z lines outside the occupied-line mask (all are stored
// without one).
__device__ __forceinline__ bool t1_line_held(const LinePass& lp, int ix, int e) {
  return lp.lmask == nullptr || lp.lmask[ix * lp.nyz_y + e] != 0;
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

// 12-point DFT as 3 x 4: DFT4 over the stride-3 subsequences x[n1 + 3 n2],
// twiddles W12^(n1 k2), DFT3 across; X[k2 + 4 k1].
template <bool INV, typename V> struct Dft<12, INV, V> {
  __device__ __forceinline__ static void run(V* a) {
    using T = typename RealOf<V>::T;
    const T s3 = (T)0.86602540378443864676, h = (T)0.5;
    V u[3][4];
#pragma unroll
    for (int n1 = 0; n1 < 3; ++n1) {
      V t[4] = {a[n1], a[n1 + 3], a[n1 + 6], a[n1 + 9]};
      Dft<4, INV, V>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) u[n1][k2] = t[k2];
    }
    // W12^e = cos(pi e/6) -+ i sin(pi e/6)
    auto tw = [&](V v, T c, T sn) {  // v * (c - i sn) forward, v * (c + i sn) inverse
      return INV ? cx<V>(v.x * c - v.y * sn, v.y * c + v.x * sn) : cx<V>(v.x * c + v.y * sn, v.y * c - v.x * sn);
    };
    u[1][1] = tw(u[1][1], s3, h);
    u[1][2] = tw(u[1][2], h, s3);
    u[1][3] = rotq<INV>(u[1][3]);
    u[2][1] = tw(u[2][1], h, s3);
    u[2][2] = tw(u[2][2], -h, s3);
    u[2][3] = cx<V>(-u[2][3].x, -u[2][3].y);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      V t[3] = {u[0][k2], u[1][k2], u[2][k2]};
      Dft<3, INV, V>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 3; ++k1) a[k2 + 4 * k1] = t[k1];
    }
  }
};

// 16-point DFT as 4 x 4: four DFT4 over stride-4 subsequences, twiddles
// W16^(j2*k1), four DFT4 across; X[k1 + 4 k2].
template <bool INV, typename V> struct Dft<16, INV, V> {
  __device__ __forceinline__ static void run(V* a) {
    using T = typename RealOf<V>::T;
    const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, h = (T)0.70710678118654752440;
    V t[4][4];
#pragma unroll
    for (int j2 = 0; j2 < 4; ++j2) {
      V u[4] = {a[j2], a[j2 + 4], a[j2 + 8], a[j2 + 12]};
      Dft<4, INV, V>::run(u);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) t[j2][k1] = u[k1];
    }
    // W16^m = cos(2 pi m/16) -+ i sin(2 pi m/16)
    auto tw = [&](V v, T c, T s) {  // v * (c - i s) forward, v * (c + i s) inverse
      return INV ? cx<V>(v.x * c - v.y * s, v.y * c + v.x * s) : cx<V>(v.x * c + v.y * s, v.y * c - v.x * s);
    };
    t[1][1] = tw(t[1][1], c1, s1);
    t[1][2] = tw(t[1][2], h, h);
    t[1][3] = tw(t[1][3], s1, c1);
    t[2][1] = tw(t[2][1], h, h);
    t[2][2] = rotq<INV>(t[2][2]);
    t[2][3] = tw(t[2][3], -h, h);
    t[3][1] = tw(t[3][1], s1, c1);
    t[3][2] = tw(t[3][2], -h, h);
    t[3][3] = tw(t[3][3], -c1, -s1);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      V u[4] = {t[0][k1], t[1][k1], t[2][k1], t[3][k1]};
      Dft<4, INV, V>::run(u);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = u[k2];
    }
  }
};

// ------------------------------------------------------- composite radices
// cos(x) and sin(x) for |x| <= pi/2 at compile time (Taylor series, 24
// terms: the remainder is below 1e-40), for the twiddles of composite
// butterflies.
constexpr double ce_cos_q(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 24; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
constexpr double ce_sin_q(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 24; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
// cos and sin of 2 pi e / R, reduced to |x| <= pi/2 by the symmetries
// cos(pi - x) = -cos x, sin(pi - x) = sin x and the odd / even parity.
template <int R> struct TwiddleR {
  double c[R], s[R];
  constexpr TwiddleR() : c(), s() {
    constexpr double pi = 3.14159265358979323846;
    for (int e = 0; e < R; ++e) {
      // angle in [-pi, pi): x = 2 pi e' / R with e' = e or e - R
      const int ee = 2 * e < R ? e : e - R;
      double x = 2.0 * pi * ee / R;
      double sg = 1.0;
      if (x > pi / 2) {
        x = pi - x;
        sg = -1.0;
      } else if (x < -pi / 2) {
        x = -pi - x;
        sg = -1.0;
      }
      c[e] = sg * ce_cos_q(x);
      s[e] = ce_sin_q(x);
    }
  }
};

// R = A * B point DFT: DFT_B over the stride-A subsequences x[n1 + A n2],
// twiddles W_R^(n1 k2), DFT_A across; X[k2 + B k1].  Twiddles that are
// -1 or -+i are applied as sign flips and swaps.
template <int A, int B, bool INV, typename V> struct DftAB {
  __device__ __forceinline__ static void run(V* a) {
    using T = typename RealOf<V>::T;
    constexpr int R = A * B;
    constexpr TwiddleR<R> tw{};
    V u[A][B];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) {
      V t[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) t[n2] = a[n1 + A * n2];
      Dft<B, INV, V>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) {
        const int e = (n1 * k2) % R;
        const V v = t[k2];
        if (e == 0) {
          u[n1][k2] = v;
        } else if (2 * e == R) {
          u[n1][k2] = cx<V>(-v.x, -v.y);
        } else if (4 * e == R) {
          u[n1][k2] = rotq<INV>(v);
        } else if (4 * e == 3 * R) {
          u[n1][k2] = rotq<!INV>(v);
        } else {
          // v * (c - i s) forward, v * (c + i s) inverse
          const T c = (T)tw.c[e], sn = (T)tw.s[e];
          u[n1][k2] = INV ? cx<V>(v.x * c - v.y * sn, v.y * c + v.x * sn)
                          : cx<V>(v.x * c + v.y * sn, v.y * c - v.x * sn);
        }
      }
    }
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) {
      V t[A];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) t[n1] = u[n1][k2];
      Dft<A, INV, V>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) a[k2 + B * k1] = t[k1];
    }
  }
};
template <bool INV, typename V> struct Dft<6, INV, V> : DftAB<2, 3, INV, V> {};
template <bool INV, typename V> struct Dft<9, INV, V> : DftAB<3, 3, INV, V> {};
template <bool INV, typename V> struct Dft<10, INV, V> : DftAB<2, 5, INV, V> {};
template <bool INV, typename V> struct Dft<15, INV, V> : DftAB<3, 5, INV, V> {};

// Thread mapping of a tile [m][QQ] (QQ = P*Q columns): thread t owns column
// qq = t % QQ and the element rows rt, rt + TPC, rt + 2 TPC, ... with
// rt = t / QQ and TPC = blockDim / QQ threads per column (TPC = ceil(m/VPT)).
// Lanes of a wavefront therefore walk consecutive columns, i.e. consecutive
// LDS words and consecutive global addresses.
struct TileMap {
  int QQ, TPC, qq, rt;
};

// One Stockham pass of radix R: every column has m/R butterflies; this thread
// does butterflies j = rt + i*TPC.  Twiddles come from the LDS copy of the
// length-m table.  NB*TPC >= m/R because TPC = ceil(m/VPT).
template <int R, bool INV, int VPT, typename V>
__device__ __forceinline__ void stockham_pass(V* buf, const V* tws, int M, int Ns, const TileMap& tm) {
  constexpr int NB = (VPT + R - 1) / R;
  const int nb = M / R;
  const int ts = M / (Ns * R);
  const int QQ = tm.QQ;
  V v[NB][R];
  int jj[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int j = tm.rt + i * tm.TPC;
    jj[i] = j;
    if (j < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[(j + r * nb) * QQ + tm.qq];
      const int jm = j % Ns;
      if (jm != 0) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          V w = tws[jm * r * ts];
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
    const int j = jj[i];
    if (j < nb) {
      const int base = (j / Ns) * Ns * R + j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(base + r * Ns) * QQ + tm.qq] = v[i][r];
    }
  }
  __syncthreads();
}

template <bool INV, int VPT, typename V>
__device__ void lds_fft(V* buf, const V* tws, const LinePass& lp, const TileMap& tm) {
  int Ns = 1;
  for (int p = 0; p < lp.nrad; ++p) {
    const int R = lp.rad[p];
    switch (R) {
      case 8: stockham_pass<8, INV, VPT>(buf, tws, lp.m, Ns, tm); break;
      case 4: stockham_pass<4, INV, VPT>(buf, tws, lp.m, Ns, tm); break;
      case 2: stockham_pass<2, INV, VPT>(buf, tws, lp.m, Ns, tm); break;
      case 3: stockham_pass<3, INV, VPT>(buf, tws, lp.m, Ns, tm); break;
      default: stockham_pass<5, INV, VPT>(buf, tws, lp.m, Ns, tm); break;
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
    const V* g = G + (((long long)ay * lp.hz + az) * lp.hx + ax) * 6;
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
  using BV = typename BoundaryOf<V, MODE>::T;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int M = lp.m, P = lp.P, Q = lp.Q, QQ = P * Q;
  V* buf = reinterpret_cast<V*>(smem_raw);
  V* tws = buf + M * QQ;
  const int tid = threadIdx.x, nthr = blockDim.x;
  TileMap tm;
  tm.QQ = QQ;
  tm.TPC = nthr / QQ;
  tm.qq = tid % QQ;
  tm.rt = tid / QQ;
  const int p = tm.qq / Q, q = tm.qq % Q;
  const int pt0 = blockIdx.x * P;
  const int lo = blockIdx.y;
  const int np_here = min(P, lp.npts - pt0);
  const bool active = tm.rt < tm.TPC && p < np_here;
  const int pt = pt0 + p;
  const int zl = zline_of<MODE>(lp, pt, active);  // the point's line (z passes)
  const V* __restrict__ src = reinterpret_cast<const V*>(lp.src) + (long long)lo * lp.lo_src;

  // ---- load: every global load of this thread (inputs, then twiddles) is
  // issued before any LDS store
  V v[VPT];
  if (is_scatter<MODE>() || MODE == PASS_CUBOID_Z) {
    const int c = q / lp.Lb, l = lp.col0 + lo * lp.Lb + q % lp.Lb;
    long long mrow[VPT];
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int e = tm.rt + i * tm.TPC;
      mrow[i] = -1;
      if (active && e < lp.nin && l < lp.L) {
        const long long cell = zline_cell(lp, zl) + e;
        mrow[i] = is_scatter<MODE>() ? (long long)lp.inv[cell] : cell;
      }
    }
    const BV* __restrict__ S = reinterpret_cast<const BV*>(lp.S);
#pragma unroll
    for (int i = 0; i < VPT; ++i) v[i] = cvt<V>(load_or_zero(mrow[i] >= 0, S + (3 * mrow[i] + c) * lp.L + l));
  } else {
    const int a = zl / lp.ptb, b = zl % lp.ptb;
    const V* __restrict__ base = src + a * lp.sa + b * lp.sb + q;
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int e = tm.rt + i * tm.TPC;
      v[i] = load_or_zero(active && e < lp.nin, base + e * lp.se);
    }
  }
  const V twv = load_or_zero(tid < M, reinterpret_cast<const V*>(lp.tw) + tid);
  if (tid < M) tws[tid] = twv;
  for (int i = tid + nthr; i < M; i += nthr) tws[i] = reinterpret_cast<const V*>(lp.tw)[i];
#pragma unroll
  for (int i = 0; i < VPT; ++i) {
    const int e = tm.rt + i * tm.TPC;
    if (tm.rt < tm.TPC && e < M) buf[e * QQ + tm.qq] = v[i];
  }
  __syncthreads();

  // ---- transform ------------------------------------------------------------
  if (is_gather<MODE>() || MODE == PASS_CROP_Z || (MODE == PASS_PLAIN && lp.inverse)) {
    lds_fft<true, VPT>(buf, tws, lp, tm);
  } else {
    lds_fft<false, VPT>(buf, tws, lp, tm);
    if (MODE == PASS_CONV_X) {
      green_multiply(buf, lp, pt0, np_here, QQ, tid, nthr);
      lds_fft<true, VPT>(buf, tws, lp, tm);
    }
  }

  // ---- store (cropped to nout) ----------------------------------------------
  if (is_gather<MODE>() || MODE == PASS_CROP_Z) {
    using T = typename RealOf<V>::T;
    using TB = typename RealOf<BV>::T;
    const T sc = (T)lp.scale;
    const int c = q / lp.Lb, l = lp.col0 + lo * lp.Lb + q % lp.Lb;
    if (!active || l >= lp.L) return;
    V* __restrict__ out = reinterpret_cast<V*>(lp.AS);
    long long row[VPT];
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int e = tm.rt + i * tm.TPC;
      row[i] = -1;
      if (e < lp.nout) {
        const long long cell = zline_cell(lp, zl) + e;
        const long long m = is_gather<MODE>() ? (long long)lp.inv[cell] : cell;
        if (m >= 0) row[i] = 3 * m + c;
      }
    }
    if (is_gather<MODE>()) {
      const BV* __restrict__ S = reinterpret_cast<const BV*>(lp.S);
      const BV* __restrict__ dg = reinterpret_cast<const BV*>(lp.diag);
      const BV* __restrict__ jp = reinterpret_cast<const BV*>(lp.jpre);
      BV* __restrict__ outb = reinterpret_cast<BV*>(lp.AS);
      const TB scb = (TB)lp.scale;
      BV sv[VPT], dv[VPT], jv[VPT];
#pragma unroll
      for (int i = 0; i < VPT; ++i) {
        if (row[i] >= 0) {
          sv[i] = S[row[i] * lp.L + l];
          dv[i] = dg[row[i]];
          jv[i] = jp[row[i]];
        }
      }
#pragma unroll
      for (int i = 0; i < VPT; ++i) {
        if (row[i] >= 0) {
          const int e = tm.rt + i * tm.TPC;
          const BV y = cscale(cvt<BV>(buf[e * QQ + tm.qq]), scb);
          outb[row[i] * lp.L + l] = cmul(jv[i], cadd(cmul(dv[i], sv[i]), y));
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < VPT; ++i) {
        if (row[i] >= 0) {
          const int e = tm.rt + i * tm.TPC;
          out[row[i] * lp.L + l] = cscale(buf[e * QQ + tm.qq], sc);
        }
      }
    }
    return;
  }
  if (!active) return;
  V* __restrict__ dst = reinterpret_cast<V*>(lp.dst) + (long long)lo * lp.lo_dst;
  const int a = zl / lp.ptb, b = zl % lp.ptb;
  V* __restrict__ obase = dst + a * lp.da + b * lp.db + q;
  const bool ymask = MODE == PASS_PLAIN && lp.inverse;  // y inverse writes T1
#pragma unroll
  for (int i = 0; i < VPT; ++i) {
    const int e = tm.rt + i * tm.TPC;
    if (e < lp.nout && (!ymask || t1_line_held(lp, a, e))) obase[e * lp.de] = buf[e * QQ + tm.qq];
  }
}

// ---------------------------------------------------------------------------
// Two-step register FFT for m = R1 * R2 (R1, R2 in {2,3,4,5,6,8,9,10,12,15,16}).
// Step 1: thread (qq, j2) holds x[j1*R2 + j2] (j1 < R1) in registers, does a
// DFT_R1, multiplies by W_m^(j2*k1) and writes LDS[k1*R2 + j2].  Step 2:
// thread (qq, k1) reads LDS[k1*R2 + j2] (j2 < R2), does a DFT_R2 and holds
// X[k1 + R1*k2].  One LDS exchange per transform; the inputs come straight
// from global memory (or LDS) and the outputs go straight to global memory
// (or LDS), so a pass costs two LDS trips instead of 2 per radix.
template <int R1, int R2, bool INV, typename V>
__device__ __forceinline__ void reg2_step1(V* a, const V* tws, int j2) {
  Dft<R1, INV, V>::run(a);
  if (j2 != 0) {
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) {
      V w = tws[j2 * k1];
      if (INV) w.y = -w.y;
      a[k1] = cmul(a[k1], w);
    }
  }
}

// Entry i (of n) of the x pass's staged Green lines, laid out [p][qx][6] for
// the points pt0, pt0 + 1, ... (zero for i >= n).  The half spectrum is stored
// x-innermost, [(ay*hz + az)*hx + ax][6]; the odd components (xy, xz, yz)
// flip sign under the mirror q -> m - q.
constexpr int kGreenPerThread = 4;
template <typename V, int M>
__device__ __forceinline__ V green_entry(const LinePass& lp, int pt0, int i, int n) {
  const int comp = i % 6, qx = (i / 6) % M, pp = i / (6 * M);
  const int ptg = pt0 + pp;
  const int ky = ptg / lp.mz, kz = ptg - ky * lp.mz;
  int ax = qx, ay = ky, az = kz;
  bool sgn = false;
  if (ax > lp.mx - ax) { ax = lp.mx - ax; sgn ^= (comp == 1 || comp == 2); }
  if (ay > lp.my - ay) { ay = lp.my - ay; sgn ^= (comp == 1 || comp == 4); }
  if (az > lp.mz - az) { az = lp.mz - az; sgn ^= (comp == 2 || comp == 4); }
  const V* __restrict__ G = reinterpret_cast<const V*>(lp.green);
  const V g = load_or_zero(i < n, G + (((long long)ay * lp.hz + az) * lp.hx + ax) * 6 + comp);
  return sgn ? cx<V>(-g.x, -g.y) : g;
}

// Points per workgroup of a register kernel on a compiled tile of Q columns
// per point: as many as the thread bound and a 72 KB tile allow, at least
// one.  Points are consecutive along the innermost axis, so a block's lines
// are adjacent in memory.  One point's chunk per element is Q*sizeof(V)
// bytes: 384 B (three cache lines) for a complex128 8-column tile, but
// 192 B (1.5 lines) for complex64 or a 4-column tile, so those take two
// points (complex64 blocks on a compiled tile go up to 1024 threads for
// that; runtime tiles keep 512 and the register budget that comes with it).
// The z passes' lines are contiguous already (a z line is one run of
// memory), so they keep 512-thread blocks, which measured faster there.
// Waves per SIMD the register kernels are compiled for: the complex128 z
// passes (512-thread blocks, two waves per SIMD each) are held to 128 VGPRs
// so that two blocks share a CU, on the lengths that fit that without
// spilling (second radix up to 12).
template <typename V, int R2, int MODE>
constexpr int reg2_waves_per_simd() {
  return (sizeof(V) == sizeof(double2) && is_z_pass<MODE>() && R2 <= 12) ? 4 : 2;
}
template <typename V, int LB, int MODE>
constexpr int reg2_max_threads() {
  return sizeof(V) == sizeof(float2) && LB > 0 && !is_z_pass<MODE>() ? 1024 : 512;
}
constexpr int kRegTileBytes = 4608 * 16;
template <typename V, int MODE>
constexpr int reg2_points(int M, int RMAX, int Q) {
  const int by_threads = reg2_max_threads<V, 1, MODE>() / (Q * RMAX);
  const int by_lds = (kRegTileBytes / (int)sizeof(V)) / (M * Q);
  const int p = by_threads < by_lds ? by_threads : by_lds;
  return p < 1 ? 1 : p;
}

// LB > 0: column tile fixed at LB (Q = 3 LB columns per point, P points per
// workgroup from reg2_points), so that the tile width and every LDS offset
// are compile-time constants; LB = 0: tile taken from the launch arguments.
template <typename V, int R1, int R2, int MODE, int LB>
__global__ void __launch_bounds__((reg2_max_threads<V, LB, MODE>()))
__attribute__((amdgpu_waves_per_eu(reg2_waves_per_simd<V, R2, MODE>()))) k_reg2(LinePass lp) {
  constexpr int M = R1 * R2;
  constexpr int RMAX = R1 > R2 ? R1 : R2;
  constexpr bool INV_PASS = (is_gather<MODE>() || MODE == PASS_CROP_Z || MODE == PASS_PLAIN_INV);
  using BV = typename BoundaryOf<V, MODE>::T;
  constexpr int PB = LB ? reg2_points<V, MODE>(M, RMAX, 3 * LB) : 0;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Lb = LB ? LB : lp.Lb;
  const int Q = LB ? 3 * LB : lp.Q;
  const int P = LB ? PB : lp.P;
  const int QQ = P * Q;
  const int nthr = LB ? PB * 3 * LB * RMAX : (int)blockDim.x;
  V* buf = reinterpret_cast<V*>(smem_raw);
  V* tws = buf + M * QQ;
  V* gls = tws + M;  // Green tensor of the tile's points: [p][qx][6]
  const int tid = threadIdx.x;
  const int qq = tid % QQ, w = tid / QQ;
  const int p = qq / Q, q = qq % Q;
  // x pass: 1-D grid remapped so that the column tiles of one point run back
  // to back on one XCD (blocks are dealt round-robin to the 8 XCDs), where
  // the point's Green tile then stays in that XCD's L2.
  int pblk = blockIdx.x, lo_blk = blockIdx.y;
  if (MODE == PASS_CONV_X) {
    const int b = blockIdx.x, xcd = b & 7, slot = b >> 3;
    pblk = (slot / lp.nlo) * 8 + xcd;
    lo_blk = slot % lp.nlo;
    if (pblk * P >= lp.npts) return;
  }
  const int pt0 = pblk * P;
  const int np_here = min(P, lp.npts - pt0);
  const bool active = p < np_here;
  const int pt = pt0 + p;
  const int zl = zline_of<MODE>(lp, pt, active);  // the point's line (z passes)
  const bool s1 = w < R2, s2 = w < R1;
  const int j2 = w, k1 = w;
  {
    const int lo = lo_blk;
    // ---- global loads, all issued before the first LDS store so that their
    // latencies overlap: the point's inputs, the x pass's Green lines, and
    // the twiddle table ----------------------------------------------------------
    V a[RMAX];
    const V* __restrict__ src = reinterpret_cast<const V*>(lp.src) + (long long)lo * lp.lo_src;
    if (is_scatter<MODE>() || MODE == PASS_CUBOID_Z) {
      const int c = q / Lb, l = lp.col0 + lo * Lb + q % Lb;
      long long mrow[R1];
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) {
        const int e = j1 * R2 + j2;
        mrow[j1] = -1;
        if (s1 && active && e < lp.nin && l < lp.L) {
          const long long cell = zline_cell(lp, zl) + e;
          mrow[j1] = is_scatter<MODE>() ? (long long)lp.inv[cell] : cell;
        }
      }
      const BV* __restrict__ S = reinterpret_cast<const BV*>(lp.S);
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1)
        a[j1] = cvt<V>(load_or_zero(mrow[j1] >= 0, S + (3 * mrow[j1] + c) * lp.L + l));
    } else {
      const int ga = zl / lp.ptb, gb = zl % lp.ptb;
      const int se = (int)lp.se;  // a line spans fewer than 2^31 elements (checked at launch)
      const V* __restrict__ base = src + ga * lp.sa + gb * lp.sb + q + j2 * se;
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) {
        const int e = j1 * R2 + j2;
        a[j1] = load_or_zero(s1 && active && e < lp.nin, base + j1 * (R2 * se));
      }
    }
    // Green lines of the tile's points (x pass), kGreenPerThread per thread in
    // registers; the launcher keeps np * M * 6 within kGreenPerThread * nthr
    const int ng = MODE == PASS_CONV_X ? np_here * M * 6 : 0;
    V gv[kGreenPerThread];
    if (MODE == PASS_CONV_X) {
#pragma unroll
      for (int u = 0; u < kGreenPerThread; ++u) gv[u] = green_entry<V, M>(lp, pt0, tid + u * nthr, ng);
    }
    // z-inverse with gather: occupied index of each output cell this thread
    // stores (-1 for empty cells and cropped elements)
    int outm[is_gather<MODE>() ? R2 : 1];
    if (is_gather<MODE>()) {
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) {
        const int e = k1 + R1 * k2;
        const bool ok = s2 && active && e < lp.nout;
        outm[k2] = ok ? lp.inv[zline_cell(lp, zl) + e] : -1;
      }
    }
    const V twv = load_or_zero(tid < M, reinterpret_cast<const V*>(lp.tw) + tid);
    if (tid < M) tws[tid] = twv;
    for (int i = tid + nthr; i < M; i += nthr) tws[i] = reinterpret_cast<const V*>(lp.tw)[i];
    if (MODE == PASS_CONV_X) {
#pragma unroll
      for (int u = 0; u < kGreenPerThread; ++u)
        if (tid + u * nthr < ng) gls[tid + u * nthr] = gv[u];
      for (int i = tid + kGreenPerThread * nthr; i < ng; i += nthr) gls[i] = green_entry<V, M>(lp, pt0, i, ng);
    }
    __syncthreads();  // staged tables ready
    if (s1) {
      reg2_step1<R1, R2, INV_PASS>(a, tws, j2);
#pragma unroll
      for (int kk = 0; kk < R1; ++kk) buf[(kk * R2 + j2) * QQ + qq] = a[kk];
    }
    __syncthreads();
    // ---- step 2 ---------------------------------------------------------------
    if (s2) {
#pragma unroll
      for (int jj = 0; jj < R2; ++jj) a[jj] = buf[(k1 * R2 + jj) * QQ + qq];
      Dft<R2, INV_PASS, V>::run(a);
    }
    if (MODE == PASS_CONV_X) {
      // X[k1 + R1 k2] -> LDS natural order, 3x3 multiply, inverse transform
      __syncthreads();
      if (s2) {
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) buf[(k1 + R1 * k2) * QQ + qq] = a[k2];
      }
      __syncthreads();
      for (int t = tid; t < np_here * Lb * M; t += nthr) {
        const int li = t % Lb, kx = (t / Lb) % M, pp = t / (Lb * M);
        const V* g = gls + (pp * M + kx) * 6;
        V* row = buf + kx * QQ + pp * Q;
        const V x0 = row[li], x1 = row[Lb + li], x2 = row[2 * Lb + li];
        row[li] = cadd(cadd(cmul(g[0], x0), cmul(g[1], x1)), cmul(g[2], x2));
        row[Lb + li] = cadd(cadd(cmul(g[1], x0), cmul(g[3], x1)), cmul(g[4], x2));
        row[2 * Lb + li] = cadd(cadd(cmul(g[2], x0), cmul(g[4], x1)), cmul(g[5], x2));
      }
      __syncthreads();
      if (s1) {
#pragma unroll
        for (int j1 = 0; j1 < R1; ++j1) a[j1] = buf[(j1 * R2 + j2) * QQ + qq];
      }
      __syncthreads();
      if (s1) {
        reg2_step1<R1, R2, true>(a, tws, j2);
#pragma unroll
        for (int kk = 0; kk < R1; ++kk) buf[(kk * R2 + j2) * QQ + qq] = a[kk];
      }
      __syncthreads();
      if (s2) {
#pragma unroll
        for (int jj = 0; jj < R2; ++jj) a[jj] = buf[(k1 * R2 + jj) * QQ + qq];
        Dft<R2, true, V>::run(a);
      }
    }
    // ---- store X[k1 + R1 k2], k < nout ---------------------------------------------
    if (s2 && active) {
      if (is_gather<MODE>() || MODE == PASS_CROP_Z) {
        using T = typename RealOf<V>::T;
        using TB = typename RealOf<BV>::T;
        const T sc = (T)lp.scale;
        const int c = q / Lb, l = lp.col0 + lo * Lb + q % Lb;
        if (l < lp.L) {
          V* __restrict__ out = reinterpret_cast<V*>(lp.AS);
          // rows 3m + c of the output cells (m from the prefetched index)
          int row[R2];
#pragma unroll
          for (int k2 = 0; k2 < R2; ++k2) {
            const int e = k1 + R1 * k2;
            const int m = is_gather<MODE>() ? outm[k2] : (e < lp.nout ? (int)zline_cell(lp, zl) + e : -1);
            row[k2] = m >= 0 ? 3 * m + c : -1;
          }
          if (is_gather<MODE>()) {
            const BV* __restrict__ S = reinterpret_cast<const BV*>(lp.S);
            const BV* __restrict__ dg = reinterpret_cast<const BV*>(lp.diag);
            const BV* __restrict__ jp = reinterpret_cast<const BV*>(lp.jpre);
            BV* __restrict__ outb = reinterpret_cast<BV*>(lp.AS);
            const TB scb = (TB)lp.scale;
            // groups of four outputs keep the staged operands within budget
            constexpr int H = R2 < 4 ? R2 : 4;
#pragma unroll
            for (int k0 = 0; k0 < R2; k0 += H) {
              BV sv[H], dv[H], jv[H];
#pragma unroll
              for (int h = 0; h < H && k0 + h < R2; ++h) {
                const int r = row[k0 + h];
                sv[h] = load_or_zero(r >= 0, S + (long long)r * lp.L + l);
                dv[h] = load_or_zero(r >= 0, dg + r);
                jv[h] = load_or_zero(r >= 0, jp + r);
              }
#pragma unroll
              for (int h = 0; h < H && k0 + h < R2; ++h) {
                const int r = row[k0 + h];
                if (r >= 0)
                  outb[(long long)r * lp.L + l] =
                      cmul(jv[h], cadd(cmul(dv[h], sv[h]), cscale(cvt<BV>(a[k0 + h]), scb)));
              }
            }
          } else {
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2)
              if (row[k2] >= 0) out[(long long)row[k2] * lp.L + l] = cscale(a[k2], sc);
          }
        }
      } else {
        V* __restrict__ dst = reinterpret_cast<V*>(lp.dst) + (long long)lo * lp.lo_dst;
        const int ga = zl / lp.ptb, gb = zl % lp.ptb;
        const int de = (int)lp.de;
        V* __restrict__ obase = dst + ga * lp.da + gb * lp.db + q + k1 * de;
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) {
          const int e = k1 + R1 * k2;
          const bool held = MODE != PASS_PLAIN_INV || t1_line_held(lp, ga, e);  // y inverse writes T1
          if (e < lp.nout && held) obase[k2 * (R1 * de)] = a[k2];
        }
      }
    }
  }
}

// Inverse 16-point DFT, first half of the outputs only:
// x[n] = sum_k y[k] W16^(-n k) for n < 8, with y[k] = col[k * stride] (LDS).
// Radix 4 x 4 with k = b + 4 j: for each b, a DFT4 over j gives u_b[kk];
// x[kk + 4 m] = sum_b W16^(-b kk) u_b[kk] W4^(-b m), and the b sum (m = 0, 1)
// is accumulated as each group of four inputs arrives, so eight accumulators
// and one group stay live instead of all sixteen values and their partials.
template <typename V>
__device__ __forceinline__ void idft16_lower_half(const V* col, int stride, V* x) {
  using T = typename RealOf<V>::T;
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, h = (T)0.70710678118654752440;
  // v * exp(+2 pi i e / 16) for e = b * kk (inverse twiddle)
  auto tw = [&](V v, int e) -> V {
    switch (e) {
      case 1: return cx<V>(v.x * c1 - v.y * s1, v.y * c1 + v.x * s1);
      case 2: return cx<V>((v.x - v.y) * h, (v.y + v.x) * h);
      case 3: return cx<V>(v.x * s1 - v.y * c1, v.y * s1 + v.x * c1);
      case 4: return cx<V>(-v.y, v.x);
      case 6: return cx<V>((-v.x - v.y) * h, (v.x - v.y) * h);
      case 9: return cx<V>(-v.x * c1 + v.y * s1, -v.y * c1 - v.x * s1);
      default: return v;
    }
  };
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    V u[4] = {col[b * stride], col[(b + 4) * stride], col[(b + 8) * stride], col[(b + 12) * stride]};
    Dft<4, true, V>::run(u);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const V v = tw(u[kk], b * kk);
      // W4^(-b) = i^b
      const V r = b == 0 ? v : b == 1 ? cx<V>(-v.y, v.x) : b == 2 ? cx<V>(-v.x, -v.y) : cx<V>(v.y, -v.x);
      x[kk] = b == 0 ? v : cadd(x[kk], v);
      x[kk + 4] = b == 0 ? r : cadd(x[kk + 4], r);
    }
  }
}

// The six Green entries of wavenumber (qx, ky, kz) of the x pass's point pt
// ((ky, kz), kz fastest), mirrored from the stored half spectrum.
template <typename V>
__device__ __forceinline__ void green_row(const LinePass& lp, int pt, int qx, V* g) {
  const int ky = pt / lp.mz, kz = pt - ky * lp.mz;
  int ax = qx, ay = ky, az = kz;
  bool fx = false, fy = false, fz = false;
  if (ax > lp.mx - ax) { ax = lp.mx - ax; fx = true; }
  if (ay > lp.my - ay) { ay = lp.my - ay; fy = true; }
  if (az > lp.mz - az) { az = lp.mz - az; fz = true; }
  const V* __restrict__ G = reinterpret_cast<const V*>(lp.green) + (((long long)ay * lp.hz + az) * lp.hx + ax) * 6;
#pragma unroll
  for (int c = 0; c < 6; ++c) g[c] = G[c];
  if (fx != fy) g[1] = cx<V>(-g[1].x, -g[1].y);  // xy
  if (fx != fz) g[2] = cx<V>(-g[2].x, -g[2].y);  // xz
  if (fy != fz) g[4] = cx<V>(-g[4].x, -g[4].y);  // yz
}

// ---------------------------------------------------------------------------
// x pass (forward x FFT, 3x3 Green multiply, inverse x FFT) with one point
// (ky, kz) per workgroup and a compile-time tile of LB orientations: Q = 3 LB
// columns, component-major; m = R1 * R2.
//  forward   thread (q, j2 < R2) holds x[j1 R2 + j2] (j1 < R1): DFT_R1 and
//            twiddle W^(j2 k1) into LDS row k1 R2 + j2; thread (q, k1 < R1)
//            reads rows k1 R2 + j2 (j2 < R2): DFT_R2 gives X[k1 + R1 k2],
//            written back to the rows it read, so row k1 R2 + k2 holds
//            wavenumber k1 + R1 k2.
//  multiply  thread (row r, group g) keeps the row's six Green entries in
//            registers and forms y_c = sum_c' G_cc' x_c' for orientations
//            g, g + G, ... (G = floor(threads / m) groups; the remaining
//            threads sit this step out).
//  inverse   decimation in time from that digit-reversed order: thread
//            (q, k1) reads its rows back, inverse DFT_R2 over k2 gives
//            Z[k1][n2], twiddle W^(-n2 k1), written to the same rows; thread
//            (q, n2 < R2) reads rows k1 R2 + n2 (k1 < R1): inverse DFT_R1
//            gives x[n2 + R2 n1].
// Each thread's step-2 and inverse-step-1 writes go to rows only it reads,
// so one barrier per exchange suffices (five in all, no reorder pass).  Rows
// are S = Q | 1 values apart: an odd stride puts the multiply's lanes, which
// walk rows, on distinct LDS banks.  The launcher guarantees n_in, n_out <=
// m / 2 (every Goodman line: m >= 2n - 1 and even), so inputs j1 >= R1 / 2
// are zero and outputs n1 >= R1 / 2 are cropped: they are neither loaded nor
// stored, the zero operands fold away (this file is built without signed
// zeros) and the unused outputs are dead code.
// Inputs of one x line: thread (q, w < R2) loads x[j1 R2 + w] for the
// j1 < R1 / 2 that can be nonzero (zero beyond n_in).
template <typename V, int R1, int R2, int LB>
__device__ __forceinline__ void xconv_load(const LinePass& lp, int pt, int lo, int q, int w, V* an) {
  constexpr int J1 = R1 / 2;
  const int ga = pt / lp.ptb, gb = pt % lp.ptb;
  const int se = (int)lp.se;  // checked at launch
  const V* __restrict__ base =
      reinterpret_cast<const V*>(lp.src) + (long long)lo * lp.lo_src + ga * lp.sa + gb * lp.sb + q + w * se;
  // whole line in range (the common case): plain loads, no per-element
  // guards to keep alive until the stores
  if (w < R2 && (J1 - 1) * R2 + w < lp.nin) {
#pragma unroll
    for (int j1 = 0; j1 < J1; ++j1) an[j1] = base[j1 * (R2 * se)];
  } else {
#pragma unroll
    for (int j1 = 0; j1 < J1; ++j1) an[j1] = load_or_zero(w < R2 && j1 * R2 + w < lp.nin, base + j1 * (R2 * se));
  }
}

// Waves per SIMD the x pass is compiled for.  Its blocks are 6 waves, placed
// 2-2-1-1 over a CU's four SIMDs, so at 5 waves per SIMD (81-96 VGPRs) a
// third block rarely finds room; the complex128 m = 128 kernel fits 80
// VGPRs (6 waves per SIMD) without spilling, and then three blocks share a
// CU.  Other instances keep the compiler's default.
template <typename V, int R1, int R2, int LB>
constexpr int xconv_waves_per_simd() {
  return (sizeof(V) == sizeof(double2) && R1 * R2 == 128 && LB == 8) ? 6 : 2;
}

// Forward step 1 of the x pass on a half-zero input x[j1 R2 + w], j1 < R1/2,
// as two half-length DFTs: X[2k] = DFT_{R1/2}(x)[k] and X[2k + 1] =
// DFT_{R1/2}(x_j W_R1^j)[k].  The inputs wait in this thread's odd-k1 rows
// while the even half is formed, so only R1/2 values are live at a time
// (this step sets the kernel's register peak); each output is twiddled by
// W_M^(w k1) and written to row k1 R2 + w.
template <typename V, int R1, int R2, int S>
__device__ __forceinline__ void xconv_step1_even(V* a, V* __restrict__ buf, const V* __restrict__ tws, int w, int q) {
  constexpr int H = R1 / 2;
  constexpr int GRP = 4;  // twiddle loads in flight per group
  V* __restrict__ col = buf + w * S + q;
#pragma unroll
  for (int j = 0; j < H; ++j) col[(2 * j + 1) * R2 * S] = a[j];
  Dft<H, false, V>::run(a);
  col[0] = a[0];
#pragma unroll
  for (int k0 = 1; k0 < H; k0 += GRP) {
    V t[GRP];
#pragma unroll
    for (int k = k0; k < k0 + GRP && k < H; ++k) t[k - k0] = tws[2 * k * w];
#pragma unroll
    for (int k = k0; k < k0 + GRP && k < H; ++k) col[2 * k * R2 * S] = cmul(a[k], t[k - k0]);
    __builtin_amdgcn_sched_barrier(0);
  }
}
// The odd half: x_j W_R1^j (compile-time twiddles), DFT, twiddle, store.
template <typename V, int R1, int R2, int S>
__device__ __forceinline__ void xconv_step1_odd(V* a, V* __restrict__ buf, const V* __restrict__ tws, int w, int q) {
  using T = typename RealOf<V>::T;
  constexpr int H = R1 / 2;
  constexpr int GRP = 4;
  constexpr TwiddleR<R1> w1{};
  V* __restrict__ col = buf + w * S + q;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const V x = col[(2 * j + 1) * R2 * S];
    const T c = (T)w1.c[j], sn = (T)w1.s[j];
    a[j] = j == 0 ? x : cx<V>(x.x * c + x.y * sn, x.y * c - x.x * sn);
  }
  Dft<H, false, V>::run(a);
#pragma unroll
  for (int k0 = 0; k0 < H; k0 += GRP) {
    V t[GRP];
#pragma unroll
    for (int k = k0; k < k0 + GRP && k < H; ++k) t[k - k0] = tws[(2 * k + 1) * w];
#pragma unroll
    for (int k = k0; k < k0 + GRP && k < H; ++k) col[(2 * k + 1) * R2 * S] = cmul(a[k], t[k - k0]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// One x line tile after its inputs are in a[0, R1/2): forward transform, 3x3
// Green multiply, inverse transform, store.  Needs the twiddles staged in
// tws; buf is free again only after a barrier.
template <typename V, int R1, int R2, int LB>
__device__ __forceinline__ void xconv_point(const LinePass& lp, V* buf, const V* tws, V* a, int pt, int lo,
                                            int tid) {
  constexpr int M = R1 * R2;
  constexpr int RMAX = R1 > R2 ? R1 : R2;
  constexpr int Q = 3 * LB;
  constexpr int S = Q | 1;
  constexpr int NTHR = Q * RMAX;
  constexpr int G = NTHR / M;  // orientation groups of the multiply
  constexpr int J1 = R1 / 2;
  const int q = tid % Q, w = tid / Q;
  const bool s1 = w < R2, s2 = w < R1;
  // Green entries of this thread's multiply row, in flight during step 2
  const int r = tid % M, grp = tid / M;
  V gr[6];
  if constexpr (xconv_waves_per_simd<V, R1, R2, LB>() > 2) {
    // register-lean form: split step 1; the Green loads are issued only
    // after it (not hoisted into it), in flight over the barrier
    if (s1) {
      xconv_step1_even<V, R1, R2, S>(a, buf, tws, w, q);
      xconv_step1_odd<V, R1, R2, S>(a, buf, tws, w, q);
    }
    __builtin_amdgcn_sched_barrier(0);
    green_row<V>(lp, pt, r / R2 + R1 * (r % R2), gr);
    __syncthreads();
  } else {
#pragma unroll
    for (int j1 = J1; j1 < R1; ++j1) a[j1] = cx<V>(0, 0);
    if (s1) {
      reg2_step1<R1, R2, false>(a, tws, w);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) buf[(k1 * R2 + w) * S + q] = a[k1];
    }
    green_row<V>(lp, pt, r / R2 + R1 * (r % R2), gr);
    __syncthreads();
  }
  if (s2) {
#pragma unroll
    for (int j2 = 0; j2 < R2; ++j2) a[j2] = buf[(w * R2 + j2) * S + q];
    Dft<R2, false, V>::run(a);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) buf[(w * R2 + k2) * S + q] = a[k2];
  }
  __syncthreads();
  if (grp < G) {
    V* row = buf + r * S;
#pragma unroll
    for (int li = grp; li < LB; li += G) {
      const V x0 = row[li], x1 = row[LB + li], x2 = row[2 * LB + li];
      row[li] = cadd(cadd(cmul(gr[0], x0), cmul(gr[1], x1)), cmul(gr[2], x2));
      row[LB + li] = cadd(cadd(cmul(gr[1], x0), cmul(gr[3], x1)), cmul(gr[4], x2));
      row[2 * LB + li] = cadd(cadd(cmul(gr[2], x0), cmul(gr[4], x1)), cmul(gr[5], x2));
    }
  }
  __syncthreads();
  if (s2) {
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) a[k2] = buf[(w * R2 + k2) * S + q];
    Dft<R2, true, V>::run(a);
    if (w != 0) {
#pragma unroll
      for (int n2 = 1; n2 < R2; ++n2) {
        V tw = tws[n2 * w];
        tw.y = -tw.y;
        a[n2] = cmul(a[n2], tw);
      }
    }
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) buf[(w * R2 + n2) * S + q] = a[n2];
  }
  __syncthreads();
  if (s1) {
    if constexpr (R1 == 16) {
      idft16_lower_half<V>(buf + w * S + q, R2 * S, a);
    } else {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) a[k1] = buf[(k1 * R2 + w) * S + q];
      Dft<R1, true, V>::run(a);
    }
    const int ga = pt / lp.ptb, gb = pt % lp.ptb;
    const int de = (int)lp.de;  // checked at launch
    V* __restrict__ obase =
        reinterpret_cast<V*>(lp.dst) + (long long)lo * lp.lo_dst + ga * lp.da + gb * lp.db + q + w * de;
    if ((J1 - 1) * R2 + w < lp.nout) {
#pragma unroll
      for (int n1 = 0; n1 < J1; ++n1) obase[n1 * (R2 * de)] = a[n1];
    } else {
#pragma unroll
      for (int n1 = 0; n1 < J1; ++n1)
        if (w + R2 * n1 < lp.nout) obase[n1 * (R2 * de)] = a[n1];
    }
  }
}

// Block i of the x pass -> (point, column tile): blocks are dealt to the 8
// XCDs round-robin (i % 8 is the XCD), and the column tiles of one point are
// consecutive blocks on one XCD, where that point's Green rows then stay in
// L2.
__device__ __forceinline__ void xconv_item(const LinePass& lp, int item, int* pt, int* lo) {
  const int xcd = item & 7, slot = item >> 3;
  *pt = (slot / lp.nlo) * 8 + xcd;
  *lo = slot % lp.nlo;
}

template <typename V, int R1, int R2, int LB>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(xconv_waves_per_simd<V, R1, R2, LB>())))
k_xconv(LinePass lp) {
  constexpr int M = R1 * R2;
  constexpr int RMAX = R1 > R2 ? R1 : R2;
  constexpr int Q = 3 * LB;
  constexpr int S = Q | 1;
  constexpr int NTHR = Q * RMAX;
  static_assert(NTHR / M >= 1, "the multiply needs a thread per row");
  static_assert(R1 % 2 == 0, "pruning halves the R1 digit");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* buf = reinterpret_cast<V*>(smem_raw);
  V* tws = buf + M * S;
  const int tid = threadIdx.x;
  int pt, lo;
  xconv_item(lp, blockIdx.x, &pt, &lo);
  if (pt >= lp.npts) return;
  V a[RMAX];
  xconv_load<V, R1, R2, LB>(lp, pt, lo, tid % Q, tid / Q, a);
  {
    const V twv = load_or_zero(tid < M, reinterpret_cast<const V*>(lp.tw) + tid);
    if (tid < M) tws[tid] = twv;
    for (int i = tid + NTHR; i < M; i += NTHR) tws[i] = reinterpret_cast<const V*>(lp.tw)[i];
  }
  __syncthreads();  // twiddles staged
  xconv_point<V, R1, R2, LB>(lp, buf, tws, a, pt, lo, tid);
}

// ---------------------------------------------------------------------------
// x pass in two halves of the spectrum: one point (ky, kz) and a compile-time
// tile of LB orientations per block (Q = 3 LB columns), m = R1 R2 with R1
// even, HK = R1 / 2.  With x = x[R2 j1 + j2] (j1 < HK: every Goodman line is
// half zero) and k = k1 + R1 k2, the first forward step Y[k1][j2] =
// sum_j1 x W_R1^(j1 k1) splits by the parity of k1: Y[2k] = DFT_HK(x)[k],
// Y[2k+1] = DFT_HK(x_j1 W_R1^j1)[k].  The last inverse step splits the same
// way: for n1 < HK, x_out[R2 n1 + n2] = A[n1] + W_R1^(-n1) B[n1], with A and
// B the inverse DFT_HKs over k of the rows V[2k][n2] and V[2k+1][n2].  So the
// block runs the rest of the pipeline once per parity p (the HK R2 spectral
// rows k1 = 2k + p) through an LDS tile of m/2 rows instead of m:
//  step 1   thread (q, w = j2 < R2): both DFT_HKs of its HK inputs, twiddled
//           by W_m^(k1 j2); the even half to LDS row k R2 + w, the odd half
//           kept in registers for p = 1.
//  step 2   thread (q, w = k < HK): DFT_R2 over j2 of rows w R2 + j2,
//           written back to rows w R2 + k2 (spectral index k1 + R1 k2,
//           k1 = 2 w + p).
//  multiply thread (row r, group g): the row's six Green entries, y_c =
//           sum_c' G_cc' x_c' for orientations g, g + G, ...
//  inverse  thread (q, w = k < HK): inverse DFT_R2 over k2, twiddle
//           W_m^(-n2 k1), rows w R2 + n2; then thread (q, w = n2 < R2):
//           inverse DFT_HK over k of rows k R2 + w gives A (p = 0, kept) or
//           B, and the outputs A + W_R1^(-n1) B, n1 < HK, are stored.
// Half the LDS of k_xconv, so about twice the lines in flight per CU, and
// the first and last steps keep every thread busy.
template <int R1, int R2, int LB>
constexpr int xsplit_threads() {
  return 3 * LB * (R2 > R1 / 2 ? R2 : R1 / 2);
}
// Waves per SIMD the split x pass is compiled for: four (<= 128 VGPRs) keeps
// the blocks that its LDS allows resident; the complex128 instances with a
// second radix of 12 or more need more registers than that without spills.
template <typename V, int R2>
constexpr int xsplit_waves_per_simd() {
  return (sizeof(V) == sizeof(double2) && R2 >= 12) ? 2 : (sizeof(V) == sizeof(double2) && R2 >= 10 ? 3 : 4);
}
template <typename V, int R1, int R2, int LB>
__global__ void __launch_bounds__((xsplit_threads<R1, R2, LB>()))
__attribute__((amdgpu_waves_per_eu(xsplit_waves_per_simd<V, R2>()))) k_xconv_split(LinePass lp) {
  constexpr int HK = R1 / 2, M = R1 * R2, H = HK * R2, Q = 3 * LB, S = Q | 1;
  constexpr int NTHR = xsplit_threads<R1, R2, LB>();
  constexpr int G = NTHR / H;  // orientation groups of the multiply
  static_assert(R1 % 2 == 0 && G >= 1, "even first radix; a thread per multiply row");
  using T = typename RealOf<V>::T;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* buf = reinterpret_cast<V*>(smem_raw);
  V* tws = buf + H * S;
  const int tid = threadIdx.x, q = tid % Q, w = tid / Q;
  const bool s1 = w < R2, s2 = w < HK;
  int pt, lo;
  xconv_item(lp, blockIdx.x, &pt, &lo);
  if (pt >= lp.npts) return;
  V e[HK];
  xconv_load<V, R1, R2, LB>(lp, pt, lo, q, w, e);
  for (int i = tid; i < M; i += NTHR) tws[i] = reinterpret_cast<const V*>(lp.tw)[i];
  // step 1: odd half from x_j1 W_R1^j1 (compile-time twiddles), even half in place
  constexpr TwiddleR<R1> wr{};
  V o[HK];
#pragma unroll
  for (int j = 0; j < HK; ++j) {
    const T c = (T)wr.c[j], sn = (T)wr.s[j];
    o[j] = j == 0 ? e[j] : cx<V>(e[j].x * c + e[j].y * sn, e[j].y * c - e[j].x * sn);
  }
  Dft<HK, false, V>::run(e);
  Dft<HK, false, V>::run(o);
  __syncthreads();  // twiddles staged
  if (s1) {
#pragma unroll
    for (int k = 0; k < HK; ++k) {
      buf[(k * R2 + w) * S + q] = k == 0 ? e[0] : cmul(e[k], tws[2 * k * w]);
      if (w != 0) o[k] = cmul(o[k], tws[(2 * k + 1) * w]);
    }
  }
  const int r = tid % H, grp = tid / H;
  V A[HK];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    if (p == 1 && s1) {
#pragma unroll
      for (int k = 0; k < HK; ++k) buf[(k * R2 + w) * S + q] = o[k];
    }
    __syncthreads();
    // step 2
    V a[R2 > HK ? R2 : HK];
    if (s2) {
#pragma unroll
      for (int j2 = 0; j2 < R2; ++j2) a[j2] = buf[(w * R2 + j2) * S + q];
      Dft<R2, false, V>::run(a);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) buf[(w * R2 + k2) * S + q] = a[k2];
    }
    // the Green row of this thread's multiply row, in flight over the barrier
    V gr[6];
    if (grp < G) green_row<V>(lp, pt, 2 * (r / R2) + p + R1 * (r % R2), gr);
    __syncthreads();
    // multiply
    if (grp < G) {
      V* row = buf + r * S;
#pragma unroll
      for (int li = grp; li < LB; li += G) {
        const V x0 = row[li], x1 = row[LB + li], x2 = row[2 * LB + li];
        row[li] = cadd(cadd(cmul(gr[0], x0), cmul(gr[1], x1)), cmul(gr[2], x2));
        row[LB + li] = cadd(cadd(cmul(gr[1], x0), cmul(gr[3], x1)), cmul(gr[4], x2));
        row[2 * LB + li] = cadd(cadd(cmul(gr[2], x0), cmul(gr[4], x1)), cmul(gr[5], x2));
      }
    }
    __syncthreads();
    // inverse step 2 and its twiddles W_m^(-n2 k1), k1 = 2 w + p
    if (s2) {
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) a[k2] = buf[(w * R2 + k2) * S + q];
      Dft<R2, true, V>::run(a);
      const int k1 = 2 * w + p;
      if (k1 != 0) {
#pragma unroll
        for (int n2 = 1; n2 < R2; ++n2) {
          V tw = tws[n2 * k1];
          tw.y = -tw.y;
          a[n2] = cmul(a[n2], tw);
        }
      }
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) buf[(w * R2 + n2) * S + q] = a[n2];
    }
    __syncthreads();
    // inverse step 1 over k for n2 = w
    if (s1) {
#pragma unroll
      for (int k = 0; k < HK; ++k) a[k] = buf[(k * R2 + w) * S + q];
      Dft<HK, true, V>::run(a);
      if (p == 0) {
#pragma unroll
        for (int n1 = 0; n1 < HK; ++n1) A[n1] = a[n1];
      } else {
#pragma unroll
        for (int n1 = 0; n1 < HK; ++n1) {
          // A + W_R1^(-n1) B
          const T c = (T)wr.c[n1], sn = (T)wr.s[n1];
          const V b = a[n1];
          A[n1] = n1 == 0 ? cadd(A[0], b) : cx<V>(A[n1].x + b.x * c - b.y * sn, A[n1].y + b.y * c + b.x * sn);
        }
      }
    }
    if (p == 0) __syncthreads();  // the rows are read before the odd half overwrites them
  }
  if (s1) {
    const int ga = pt / lp.ptb, gb = pt % lp.ptb;
    const int de = (int)lp.de;  // checked at launch
    V* __restrict__ obase =
        reinterpret_cast<V*>(lp.dst) + (long long)lo * lp.lo_dst + ga * lp.da + gb * lp.db + q + w * de;
    if ((HK - 1) * R2 + w < lp.nout) {
#pragma unroll
      for (int n1 = 0; n1 < HK; ++n1) obase[n1 * (R2 * de)] = A[n1];
    } else {
#pragma unroll
      for (int n1 = 0; n1 < HK; ++n1)
        if (w + R2 * n1 < lp.nout) obase[n1 * (R2 * de)] = A[n1];
    }
  }
}

constexpr int kVPT = 8;
constexpr int kTile = 4096;  // complex values per workgroup tile
// FFT kernel selection (set_fft_path): every launcher reads them
extern std::atomic<bool> g_force_generic;
extern std::atomic<bool> g_xconv_off;
// split x pass: 0 where its radix pair is balanced (R2 = R1 / 2), 1 on every
// eligible length, 2 nowhere (the one-pass kernel k_xconv throughout)
extern std::atomic<int> g_xconv_split;
template <int R1, int R2>
constexpr bool xsplit_balanced() {
  return 2 * R2 == R1;
}

template <typename V, int MODE>
void launch_mode(const LinePass& lp, dim3 grid, int nthr, size_t lds, hipStream_t s) {
  static std::once_flag attr_once;
  std::call_once(attr_once, [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_line_pass<V, kVPT, MODE>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  hipLaunchKernelGGL((k_line_pass<V, kVPT, MODE>), grid, dim3(nthr), lds, s, lp);
}

template <typename V, int R1, int R2, int MODE, int LB>
void launch_reg2_mode(const LinePass& lp, dim3 grid, int nthr, size_t lds, hipStream_t s) {
  static std::once_flag attr_once;
  std::call_once(attr_once, [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_reg2<V, R1, R2, MODE, LB>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  hipLaunchKernelGGL((k_reg2<V, R1, R2, MODE, LB>), grid, dim3(nthr), lds, s, lp);
}

template <typename V, int R1, int R2, int LB>
void launch_xconv_split(const LinePass& lp, dim3 grid, hipStream_t s) {
  static std::once_flag attr_once;
  std::call_once(attr_once, [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_xconv_split<V, R1, R2, LB>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  constexpr int S = (3 * LB) | 1;
  const size_t lds = ((size_t)(R1 / 2) * R2 * S + R1 * R2) * sizeof(V);
  hipLaunchKernelGGL((k_xconv_split<V, R1, R2, LB>), grid, dim3(xsplit_threads<R1, R2, LB>()), lds, s, lp);
}

template <typename V, int R1, int R2, int LB>
void launch_xconv(const LinePass& lp, dim3 grid, int nthr, size_t lds, hipStream_t s) {
  static std::once_flag attr_once;
  std::call_once(attr_once, [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_xconv<V, R1, R2, LB>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  hipLaunchKernelGGL((k_xconv<V, R1, R2, LB>), grid, dim3(nthr), lds, s, lp);
}

// Calls f(std::integral_constant<int, MODE>{}) for the pass mode of lp.
template <typename V, typename F>
void with_reg2_mode(const LinePass& lp, F&& f) {
  const int mode = (lp.mode == PASS_PLAIN && lp.inverse) ? PASS_PLAIN_INV : lp.mode;
  switch (mode) {
    case PASS_PLAIN: f(std::integral_constant<int, PASS_PLAIN>{}); break;
    case PASS_PLAIN_INV: f(std::integral_constant<int, PASS_PLAIN_INV>{}); break;
    case PASS_SCATTER_Z: f(std::integral_constant<int, PASS_SCATTER_Z>{}); break;
    case PASS_CONV_X: f(std::integral_constant<int, PASS_CONV_X>{}); break;
    case PASS_GATHER_Z: f(std::integral_constant<int, PASS_GATHER_Z>{}); break;
    case PASS_SCATTER_Z64:
    case PASS_GATHER_Z64:
      if constexpr (std::is_same<V, float2>::value) {
        if (mode == PASS_SCATTER_Z64)
          f(std::integral_constant<int, PASS_SCATTER_Z64>{});
        else
          f(std::integral_constant<int, PASS_GATHER_Z64>{});
      } else {
        set_error("register FFT: complex128 boundary modes need a complex64 transform");
      }
      break;
    case PASS_CUBOID_Z: f(std::integral_constant<int, PASS_CUBOID_Z>{}); break;
    default: f(std::integral_constant<int, PASS_CROP_Z>{}); break;
  }
}

// One pass on a compiled tile: points per block, threads and LDS are
// compile-time constants of the mode and tile.
template <typename V, int R1, int R2, int MODE, int LB>
void launch_reg2_fixed_mode(LinePass lp, hipStream_t s) {
  constexpr int M = R1 * R2, RMAX = R1 > R2 ? R1 : R2, Q = 3 * LB;
  constexpr bool conv = MODE == PASS_CONV_X;
  constexpr int P = reg2_points<V, MODE>(M, RMAX, Q);
  constexpr int nthr = P * Q * RMAX;
  static_assert(nthr <= reg2_max_threads<V, LB, MODE>() || P == 1, "tile exceeds the block bound");
  if (nthr > reg2_max_threads<V, LB, MODE>()) {
    set_error("register FFT launch exceeds the block's thread bound");
    return;
  }
  lp.P = P;
  const int nlo = lp.nlo > 0 ? lp.nlo : 1;
  const size_t lds = ((size_t)M * P * Q + M + (conv ? (size_t)P * M * 6 : 0)) * sizeof(V);
  const int nblk = (lp.npts + P - 1) / P;
  dim3 grid(nblk, nlo);
  if (conv) grid = dim3(((nblk + 7) / 8) * 8 * nlo, 1);
  launch_reg2_mode<V, R1, R2, MODE, LB>(lp, grid, nthr, lds, s);
}

// Column tiles with compile-time specialisations: 8 columns per component
// (the tile of L >= 8 on lengths up to 170) and 4 (lengths up to 341, e.g.
// 192 on a 96^3 cuboid).  Tile width, points per block and every LDS offset
// are then constants; other tiles take them from the launch arguments.
template <typename V, int R1, int R2, int LB>
void launch_reg2_fixed(LinePass lp, hipStream_t s) {
  constexpr int M = R1 * R2, RMAX = R1 > R2 ? R1 : R2, Q = 3 * LB;
  const bool conv = lp.mode == PASS_CONV_X;
  const int nlo = lp.nlo > 0 ? lp.nlo : 1;
  // x pass: the dedicated kernel, one point per block
  if constexpr (Q * RMAX >= M && R1 % 2 == 0) {
    if (conv && !g_xconv_off && 2 * lp.nin <= M && 2 * lp.nout <= M) {
      constexpr int S = Q | 1;
      const size_t lds_x = ((size_t)M * S + M) * sizeof(V);
      const dim3 grid_x(((lp.npts + 7) / 8) * 8 * nlo, 1);
      // the plan's choice, else the process-wide one (0 where balanced, 1 everywhere, 2 nowhere)
      const int split = lp.xsplit > 0 ? (lp.xsplit == 3 ? 0 : lp.xsplit) : g_xconv_split.load();
      if (split == 1 || (split == 0 && xsplit_balanced<R1, R2>())) {
        launch_xconv_split<V, R1, R2, LB>(lp, grid_x, s);
        return;
      }
      launch_xconv<V, R1, R2, LB>(lp, grid_x, Q * RMAX, lds_x, s);
      return;
    }
  }
  with_reg2_mode<V>(lp, [&](auto mc) { launch_reg2_fixed_mode<V, R1, R2, decltype(mc)::value, LB>(lp, s); });
}

template <typename V, int R1, int R2>
void launch_reg2(LinePass lp, hipStream_t s) {
  constexpr int M = R1 * R2, RMAX = R1 > R2 ? R1 : R2;
  // element offsets along a line are 32-bit in the kernels
  if ((long long)M * (lp.se > lp.de ? lp.se : lp.de) >= (1LL << 31)) {
    set_error("register FFT line stride exceeds 32-bit element offsets");
    return;
  }
  if (lp.Q == 3 * lp.Lb && lp.Lb == 8) return launch_reg2_fixed<V, R1, R2, 8>(lp, s);
  if (lp.Q == 3 * lp.Lb && lp.Lb == 4) return launch_reg2_fixed<V, R1, R2, 4>(lp, s);
  const bool conv = lp.mode == PASS_CONV_X;
  int P = std::min(512 / (lp.Q * RMAX), kTile / (M * lp.Q));
  // the x pass also stages each point's Green lines (M x 6 entries): on
  // narrow tiles (Q = 3) they outweigh the line tile, so the points per
  // block are also held to what fits the CU's LDS
  constexpr size_t kLdsBytes = 160 * 1024;
  if (conv) P = std::min<long long>(P, (long long)(kLdsBytes / sizeof(V) - M) / ((long long)M * (lp.Q + 6)));
  if (P < 1) P = 1;
  if (P > lp.npts) P = lp.npts;
  lp.P = P;
  const int nthr = P * lp.Q * RMAX;
  if (nthr > 512) {
    set_error("register FFT launch exceeds 512 threads");
    return;
  }
  const int nlo = lp.nlo > 0 ? lp.nlo : 1;
  const size_t lds = ((size_t)M * P * lp.Q + M + (conv ? (size_t)P * M * 6 : 0)) * sizeof(V);
  if (lds > kLdsBytes) {
    set_error("register FFT launch exceeds the LDS");
    return;
  }
  const int nblk = (lp.npts + P - 1) / P;
  dim3 grid(nblk, nlo);
  if (conv) grid = dim3(((nblk + 7) / 8) * 8 * nlo, 1);
  with_reg2_mode<V>(lp, [&](auto mc) { launch_reg2_mode<V, R1, R2, decltype(mc)::value, 0>(lp, grid, nthr, lds, s); });
}

// Register-kernel lengths m = R1 * R2, in four groups (one per translation
// unit).  R1 is even where the length allows, so the dedicated x pass and
// its pruning apply.  A column gets max(R1, R2) threads, of which R2 work in
// the first step and R1 in the second, so the pairs are as even as the
// radix set allows: 24, 36, 48, 60, 72, 80 and 96 as (8,3), (12,3), (16,3),
// (12,5), (12,6), (16,5), (16,6) ran the whole-cuboid operator 1.0-1.7x
// slower than with the pairs below (complex128, L = 64, m^3 grids).  128
// keeps (16,8) and 144 keeps (16,9): (8,16) and (12,12) measured slower.
#define BDDA_REG2_LENGTHS_A(X) X(16, 4, 4) X(30, 6, 5) X(45, 9, 5) X(60, 10, 6) X(90, 10, 9) X(128, 16, 8) X(180, 12, 15) X(256, 16, 16)
#define BDDA_REG2_LENGTHS_B(X) X(20, 4, 5) X(32, 8, 4) X(48, 8, 6) X(72, 8, 9) X(100, 10, 10) X(135, 15, 9) X(192, 16, 12)
#define BDDA_REG2_LENGTHS_C(X) X(24, 6, 4) X(36, 6, 6) X(54, 6, 9) X(75, 15, 5) X(96, 12, 8) X(120, 12, 10) X(144, 16, 9) X(225, 15, 15)
#define BDDA_REG2_LENGTHS_D(X) X(40, 8, 5) X(50, 10, 5) X(64, 8, 8) X(80, 10, 8) X(81, 9, 9) X(108, 12, 9) X(150, 10, 15) X(160, 16, 10) X(240, 16, 15)
#define BDDA_REG2_LENGTHS(X) BDDA_REG2_LENGTHS_A(X) BDDA_REG2_LENGTHS_B(X) BDDA_REG2_LENGTHS_C(X) BDDA_REG2_LENGTHS_D(X)

#define BDDA_REG2_INSTANTIATE(m, r1, r2)                                \
  template void launch_reg2<double2, r1, r2>(LinePass, hipStream_t); \
  template void launch_reg2<float2, r1, r2>(LinePass, hipStream_t);
#define BDDA_REG2_EXTERN(m, r1, r2)                                            \
  extern template void launch_reg2<double2, r1, r2>(LinePass, hipStream_t); \
  extern template void launch_reg2<float2, r1, r2>(LinePass, hipStream_t);

}  // namespace bdda
