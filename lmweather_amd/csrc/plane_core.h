// The C-grid plane's device code and its batch object PlaneCore<T> (reference cpp/plane.cpp), shared by the two
// translation units that instantiate it: plane.hip (FP64) and plane_f32.hip (FP32).  Every arithmetic block of the
// stencil and the RK4 update opens with LMW_PLANE_CONTRACT, which the including file defines: FP64 keeps
// contraction off, so each operation rounds as the reference's does and the state matches it bit for bit; FP32
// contracts a * b + c within each expression into one FMA (fewer instructions, one rounding fewer), the same way in
// the row march and in the per-stage kernel, which stay bit-identical to each other.
//
// Layout: per member h[ny][nx], u[ny][nx] (west faces), v[ny+1][nx] (south faces; rows 0 and ny are walls, zero),
// members contiguous.  The step path is k_plane_march: one launch per RK4 step, each wavefront walking north through
// a strip of 64 columns with the four stages one row apart, in registers (Bernoulli function at centres, absolute
// vorticity at corners, du, dv, dh: plane.cpp:59-129, and the stage updates: plane.cpp:143-158), so HBM sees the
// state once in and once out per step.  k_plane_stage, one launch per stage with a 64 x 4 tile per block and buffers
// x0 (state), xa, xb (stage inputs, ping-pong) and acc (RK4 sum), computes the same arithmetic in the same order and
// is kept as the reference shape for the march.
#pragma once
#ifndef LMW_PLANE_CONTRACT
#error "define LMW_PLANE_CONTRACT before including plane_core.h"
#endif
#include <algorithm>
#include <cmath>
#include <memory>

#include "common.h"
#include "expr.h"

namespace lmw {

struct PlaneGeo {
  int nx, ny, B;
  double dx, dy, g, depth, f0, beta, ly, visc, drag;
  int linear;
};

namespace {

// Field access for cell_tendency: h, u, v at (x index i, global row j) and the x neighbours of i.
// Global arrays of one member: periodic in x.
template <class T> struct GridAccess {
  const T *h, *u, *v;
  int nx;
  __device__ T H(int i, int j) const { return h[static_cast<size_t>(j) * nx + i]; }
  __device__ T U(int i, int j) const { return u[static_cast<size_t>(j) * nx + i]; }
  __device__ T V(int i, int j) const { return v[static_cast<size_t>(j) * nx + i]; }
  __device__ int xm(int i) const { return i == 0 ? nx - 1 : i - 1; }
  __device__ int xp(int i) const { return i == nx - 1 ? 0 : i + 1; }
};
// Arithmetic of the C-grid tendency on values (plane.cpp:72-128), shared by the cell-wise kernels and the row march
// so that they agree bit for bit.  Divisions by the grid spacings and their squares use the rounded reciprocal and
// one correction with the exact residual (Markstein's step): q = x r, x / d ~ q + (x - q d) r with both steps FMAs,
// three instructions where the IEEE division sequence takes ten.  It returned the correctly rounded quotient for
// every one of 2.2e9 random x over eleven divisors (1e4, 1e5, 2e4, 5e3, 3e4, 1e8, 12345.678, 7, 3, 0.1, 1e4/3);
// the only systematic difference is the sign of a zero quotient (+0 for -0 / d), which compares equal.
template <class T> struct Stencil {
  T g, dx, dy, dx2, dy2, rdx, rdy, rdx2, rdy2, drag, visc, depth;
  bool lin, viscous;
  __device__ explicit Stencil(const PlaneGeo &P) {
LMW_PLANE_CONTRACT
    g = static_cast<T>(P.g);
    dx = static_cast<T>(P.dx);
    dy = static_cast<T>(P.dy);
    dx2 = dx * dx;
    dy2 = dy * dy;
    rdx = static_cast<T>(1) / dx;
    rdy = static_cast<T>(1) / dy;
    rdx2 = static_cast<T>(1) / dx2;
    rdy2 = static_cast<T>(1) / dy2;
    drag = static_cast<T>(P.drag);
    visc = static_cast<T>(P.visc);
    depth = static_cast<T>(P.depth);
    lin = P.linear != 0;
    viscous = P.visc != 0;
  }
  __device__ static double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
  __device__ static float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
  __device__ static T quot(T x, T d, T r) {
LMW_PLANE_CONTRACT
    const T q = x * r;
    return fma_(fma_(-q, d, x), r, q);
  }
  // B = g h + K at a centre from u at its west and east faces and v at its south and north faces (plane.cpp:72-77)
  __device__ T bernoulli(T h, T uw, T ue, T vs, T vn) const {
LMW_PLANE_CONTRACT
    const T half = static_cast<T>(0.5);
    T k = 0;
    if (!lin) k = half * (half * (uw * uw + ue * ue) + half * (vs * vs + vn * vn));
    return g * h + k;
  }
  // absolute vorticity f + dv/dx - du/dy at a corner from v there and at the corner to its west and u north and south
  // of it (plane.cpp:80-92)
  __device__ T eta(T f, T v, T vw, T un, T us) const {
LMW_PLANE_CONTRACT
    T zeta = 0;
    if (!lin) {
      const T dvdx = quot(v - vw, dx, rdx);
      zeta = dvdx - quot(un - us, dy, rdy);
    }
    return f + zeta;
  }
  // du at a u point (plane.cpp:94-105): eta at the corners south (e0) and north (e1) of it, v at the corners' faces
  // west and east of them, B east (b) and west (bw) of it; ue, uw, un, us (east, west, north, south u with the wall
  // copies) enter the viscous term only
  __device__ T du(T e0, T e1, T vw0, T v0, T vw1, T v1, T b, T bw, T u, T ue, T uw, T un, T us) const {
LMW_PLANE_CONTRACT
    const T half = static_cast<T>(0.5), quarter = static_cast<T>(0.25), two = static_cast<T>(2);
    const T q = half * (e0 + e1);
    const T vbar = quarter * (vw0 + v0 + vw1 + v1);
    T r = q * vbar - quot(b - bw, dx, rdx) - drag * u;
    if (viscous) r += visc * (quot(ue - two * u + uw, dx2, rdx2) + quot(un - two * u + us, dy2, rdy2));
    return r;
  }
  // dv at a v point of rows 1..ny-1 (plane.cpp:106-117): eta at the corner there (e0) and east of it (ee0), u west and
  // east of the cells south (us, ues) and north (u, ue) of the face, B north (b) and south (bs) of it; ve, vw, vn, vs
  // enter the viscous term only
  __device__ T dv(T e0, T ee0, T us, T ues, T u, T ue, T b, T bs, T v, T ve, T vw, T vn, T vs) const {
LMW_PLANE_CONTRACT
    const T half = static_cast<T>(0.5), quarter = static_cast<T>(0.25), two = static_cast<T>(2);
    const T q = half * (e0 + ee0);
    const T ubar = quarter * (us + ues + u + ue);
    T r = -q * ubar - quot(b - bs, dy, rdy) - drag * v;
    if (viscous) r += visc * (quot(ve - two * v + vw, dx2, rdx2) + quot(vn - two * v + vs, dy2, rdy2));
    return r;
  }
  // dh by the flux form of continuity (plane.cpp:119-128): the cell's h with its west, east, north and south
  // neighbours, u at its west and east faces, v at its south and north faces; no flux through the walls
  __device__ T dh(T h, T hw, T he, T hn, T hs, T u, T ue, T v, T vn, bool has_n, bool has_s) const {
LMW_PLANE_CONTRACT
    const T half = static_cast<T>(0.5);
    const T hwf = lin ? depth : half * (hw + h);
    const T fe = (lin ? depth : half * (h + he)) * ue;
    const T fw = hwf * u;
    T fn = 0, fs = 0;
    if (has_n) fn = (lin ? depth : half * (h + hn)) * vn;
    if (has_s) fs = (lin ? depth : half * (hs + h)) * v;
    return -(quot(fe - fw, dx, rdx) + quot(fn - fs, dy, rdy));
  }
};

// f at corner row jj (plane.cpp:80-92), evaluated in double
template <class T> __device__ inline T coriolis(const PlaneGeo &P, int jj) {
LMW_PLANE_CONTRACT
  return static_cast<T>(P.f0 + P.beta * (jj * P.dy - 0.5 * P.ly));
}
// Bernoulli function at centre (ii, jj) and absolute vorticity at corner (ii, jj), jj = 0..ny, of an accessor
template <class T, class A> __device__ inline T bernoulli(const Stencil<T> &S, const A &F, int ii, int jj) {
  return S.bernoulli(F.H(ii, jj), F.U(ii, jj), F.U(F.xp(ii), jj), F.V(ii, jj), F.V(ii, jj + 1));
}
template <class T, class A> __device__ inline T abs_vorticity(const Stencil<T> &S, const PlaneGeo &P, const A &F, int ii, int jj) {
  const int ny = P.ny;
  const T un = jj < ny ? F.U(ii, jj) : F.U(ii, ny - 1), us = jj > 0 ? F.U(ii, jj - 1) : F.U(ii, 0);
  return S.eta(coriolis<T>(P, jj), F.V(ii, jj), F.V(F.xm(ii), jj), un, us);
}
// B and eta computed where they are asked for
template <class T, class A> struct OnTheFly {
  const Stencil<T> &S;
  const PlaneGeo &P;
  const A &F;
  __device__ T B(int ii, int jj) const { return bernoulli<T>(S, F, ii, jj); }
  __device__ T Eta(int ii, int jj) const { return abs_vorticity<T>(S, P, F, ii, jj); }
};

// Tendency of cell (i, j) of one member (plane.cpp:59-129); j is the global row (wall logic); E supplies the
// Bernoulli function at centres and the absolute vorticity at corners.
template <class T, class A, class E>
__device__ inline void cell_tendency(const Stencil<T> &S, const PlaneGeo &P, const A &F, const E &Q, int i, int j, T &dh, T &du, T &dv) {
  const int ny = P.ny, im = F.xm(i), ip = F.xp(i);
  const T u = F.U(i, j), v = F.V(i, j), h = F.H(i, j);
  T ue = F.U(ip, j), uw = 0, un = u, us = u;
  if (S.viscous) {
    uw = F.U(im, j);
    if (j + 1 < ny) un = F.U(i, j + 1);
    if (j > 0) us = F.U(i, j - 1);
  }
  du = S.du(Q.Eta(i, j), Q.Eta(i, j + 1), F.V(im, j), v, F.V(im, j + 1), F.V(i, j + 1), Q.B(i, j), Q.B(im, j), u, ue, uw, un, us);
  dv = 0;
  if (j >= 1) {
    T ve = 0, vw = 0, vn = 0, vs = 0;
    if (S.viscous) {
      ve = F.V(ip, j);
      vw = F.V(im, j);
      if (j + 1 < ny) vn = F.V(i, j + 1);
      if (j - 1 > 0) vs = F.V(i, j - 1);
    }
    dv = S.dv(Q.Eta(i, j), Q.Eta(ip, j), F.U(i, j - 1), F.U(ip, j - 1), u, ue, Q.B(i, j), Q.B(i, j - 1), v, ve, vw, vn, vs);
  }
  const bool has_n = j + 1 < ny, has_s = j > 0;
  dh = S.dh(h, F.H(im, j), F.H(ip, j), has_n ? F.H(i, j + 1) : h, has_s ? F.H(i, j - 1) : h, u, ue, v, has_n ? F.V(i, j + 1) : T(0), has_n, has_s);
}
template <class T, class A> __device__ inline void cell_tendency(const Stencil<T> &S, const PlaneGeo &P, const A &F, int i, int j, T &dh, T &du, T &dv) {
  cell_tendency<T>(S, P, F, OnTheFly<T, A>{S, P, F}, i, j, dh, du, dv);
}
template <class T>
__device__ inline void cell_tendency(const Stencil<T> &S, const PlaneGeo &P, const T *__restrict__ h, const T *__restrict__ u,
                                     const T *__restrict__ v, int i, int j, T &dh, T &du, T &dv) {
  cell_tendency<T>(S, P, GridAccess<T>{h, u, v, P.nx}, i, j, dh, du, dv);
}

// One RK4 stage for all members: tendency of `in`, then the stage update (plane.cpp:143-158).
template <class T>
__global__ __launch_bounds__(256) void k_plane_stage(PlaneGeo P, int stage, double dt, const T *__restrict__ in, T *__restrict__ x0,
                                                     T *__restrict__ out, T *__restrict__ acc) {
LMW_PLANE_CONTRACT
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, b = blockIdx.z;
  if (i >= P.nx || j >= P.ny) return;
  const size_t nh = static_cast<size_t>(P.nx) * P.ny, nv = static_cast<size_t>(P.nx) * (P.ny + 1), per = 2 * nh + nv;
  const T *h = in + b * per, *u = h + nh, *v = u + nh;
  T dh, du, dv;
  cell_tendency<T>(Stencil<T>(P), P, h, u, v, i, j, dh, du, dv);
  const size_t c = static_cast<size_t>(j) * P.nx + i;
  const size_t idx[3] = {b * per + c, b * per + nh + c, b * per + 2 * nh + c};
  const T k[3] = {dh, du, dv};
  const T hdt = static_cast<T>(0.5 * dt), fdt = static_cast<T>(dt), sdt = static_cast<T>(dt / 6), two = static_cast<T>(2);
  for (int f = 0; f < 3; f++) {
    const size_t q = idx[f];
    const T s0 = x0[q], kv = k[f];
    if (stage == 1) {
      acc[q] = kv;
      out[q] = s0 + hdt * kv;
    } else if (stage == 2) {
      acc[q] = acc[q] + two * kv;
      out[q] = s0 + hdt * kv;
    } else if (stage == 3) {
      acc[q] = acc[q] + two * kv;
      out[q] = s0 + fdt * kv;
    } else {
      x0[q] = s0 + sdt * (acc[q] + kv);
    }
  }
  // the northern wall row ny of v has zero tendency: carry it through the stage buffers
  if (j == P.ny - 1) {
    const size_t q = b * per + 2 * nh + static_cast<size_t>(P.ny) * P.nx + i;
    const T s0 = x0[q];
    if (stage == 1) acc[q] = 0;
    if (stage <= 3) out[q] = s0;
  }
}

// Value held by lane l - 1 (from the west) or l + 1 (from the east) of the wavefront; the end lanes read 0.  DPP
// wave_shr:1 and wave_shl:1, applied to each 32-bit half of a double.  Called only with every lane active.
__device__ inline int lane_w(int x) { return __builtin_amdgcn_mov_dpp(x, 0x138, 0xf, 0xf, true); }
__device__ inline int lane_e(int x) { return __builtin_amdgcn_mov_dpp(x, 0x130, 0xf, 0xf, true); }
__device__ inline float from_w(float x) { return __int_as_float(lane_w(__float_as_int(x))); }
__device__ inline float from_e(float x) { return __int_as_float(lane_e(__float_as_int(x))); }
__device__ inline double from_w(double x) { return __hiloint2double(lane_w(__double2hiint(x)), lane_w(__double2loint(x))); }
__device__ inline double from_e(double x) { return __hiloint2double(lane_e(__double2hiint(x)), lane_e(__double2loint(x))); }

// h and u of one row and v of its south face, per lane (column)
template <class T> struct Row {
  T h, u, v;
};
// What one RK4 stage of the row march keeps between rows besides its input rows: B at row t - 1, eta at corner row
// t, u of the east neighbour at row t - 1 and v of the west neighbour at row t.
template <class T> struct StageCarry {
  T bm, e0, uem, vw0;
};

// Advances one stage of the march by a row: m, c, p are the stage's input at rows t - 1, t, t + 1.  When
// 0 <= t < ny, k receives the tendency of row t (the arithmetic of cell_tendency, neighbours in x from the adjacent
// lanes) and the result is true.  INNER: 2 <= t <= ny - 2, where no wall condition applies.
template <bool INNER, class T>
__device__ inline bool march_row(const Stencil<T> &S, StageCarry<T> &w, const Row<T> &m, const Row<T> &c, const Row<T> &p, int t,
                                 int ny, T f1, Row<T> &k) {
  const bool has_n = INNER || t + 1 < ny, has_s = INNER || t > 0;
  // eta at corner row t + 1; north of the last row u is u(ny - 1), south of row 0 it is u(0) (plane.cpp:84-88)
  const T vw1 = from_w(p.v);
  const T e1 = S.eta(f1, p.v, vw1, has_n ? p.u : c.u, INNER || t + 1 > 0 ? c.u : p.u);
  const T ue = from_e(c.u);
  const bool live = INNER || (t >= 0 && t < ny);
  T b = 0;
  if (live) {
    b = S.bernoulli(c.h, c.u, ue, c.v, p.v);
    const T bw = from_w(b), hw = from_w(c.h), he = from_e(c.h), ee0 = from_e(w.e0);
    T uw = 0, ve = 0;
    if (S.viscous) {
      uw = from_w(c.u);
      ve = from_e(c.v);
    }
    k.u = S.du(w.e0, e1, w.vw0, c.v, vw1, p.v, b, bw, c.u, ue, uw, has_n ? p.u : c.u, has_s ? m.u : c.u);
    k.v = 0;
    if (INNER || t >= 1)
      k.v = S.dv(w.e0, ee0, m.u, w.uem, c.u, ue, b, w.bm, c.v, ve, w.vw0, has_n ? p.v : T(0), INNER || t - 1 > 0 ? m.v : T(0));
    k.h = S.dh(c.h, hw, he, has_n ? p.h : c.h, has_s ? m.h : c.h, c.u, ue, c.v, has_n ? p.v : T(0), has_n, has_s);
  }
  w.bm = b;
  w.e0 = e1;
  w.uem = ue;
  w.vw0 = vw1;
  return live;
}

// Register state of one wavefront of the row march at row r: the state at rows r - 4 .. r, the RK4 sums at rows
// r - 4 .. r - 1, the inputs of stages 2..4 at their rows t - 1 and t (stage 1 reads the state rows), the carries of
// the four stages, and the prefetched state row r + 1.
template <class T> struct March {
  Row<T> x[5], acc[4], ym[3], yc[3], next;
  StageCarry<T> w[4];
  T f[4], fnext;  // f at corner rows r - 3 .. r, and at row r + 1
};

// One row of the march: shift the windows, prefetch row r + 1, run stage s on row r - 1 - s (s = 0..3) and store
// the finished row.  INNER: every stage's row is in 2..ny-2 and row r + 1 exists, so no wall test remains.
template <bool INNER, class T>
__device__ inline void march_iteration(March<T> &M, const Stencil<T> &S, const T *__restrict__ fcor, int r, int r1, int ny, int nx,
                                       int c0, int c1, bool owner, T hdt, T fdt, T sdt, const T *__restrict__ gh, const T *__restrict__ gu,
                                       const T *__restrict__ gv, T *__restrict__ oh, T *__restrict__ ou, T *__restrict__ ov) {
LMW_PLANE_CONTRACT
  const T two = static_cast<T>(2);
#pragma unroll
  for (int i = 0; i < 4; i++) M.x[i] = M.x[i + 1];
  M.x[4] = M.next;
#pragma unroll
  for (int i = 0; i < 3; i++) M.acc[i] = M.acc[i + 1];
#pragma unroll
  for (int i = 0; i < 3; i++) M.f[i] = M.f[i + 1];
  M.f[3] = M.fnext;
  const int rn = r + 1;
  M.fnext = INNER || rn <= ny ? fcor[rn] : T(0);
  if (INNER || rn <= r1) {
    const size_t q = static_cast<size_t>(rn) * nx;
    Row<T> n{T(0), T(0), T(0)};
    if (INNER || rn < ny) {
      n.h = gh[q];
      n.u = gu[q];
    }
    if (INNER || rn <= ny) n.v = gv[q];
    M.next = n;
  }
  Row<T> p = M.x[4];  // input of stage s at row r - s
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int t = r - 1 - s;
    Row<T> k;
    const T f1 = M.f[3 - s];
    const Row<T> &m = s == 0 ? M.x[2] : M.ym[s - 1], &c = s == 0 ? M.x[3] : M.yc[s - 1];
    bool live;
    if (INNER) live = march_row<true>(S, M.w[s], m, c, p, t, ny, f1, k);
    else if (t >= 2 && t <= ny - 2) live = march_row<true>(S, M.w[s], m, c, p, t, ny, f1, k);
    else live = march_row<false>(S, M.w[s], m, c, p, t, ny, f1, k);
    if (s > 0) {
      M.ym[s - 1] = M.yc[s - 1];
      M.yc[s - 1] = p;
    }
    const Row<T> &xs = M.x[3 - s];
    Row<T> &a = M.acc[3 - s];
    if (s < 3) {
      p = xs;  // rows outside the domain and the northern wall row of v pass the state through
      if (live) {
        const T cs = s == 2 ? fdt : hdt;
        if (s == 0) a = k;
        else a = Row<T>{a.h + two * k.h, a.u + two * k.u, a.v + two * k.v};
        p = Row<T>{xs.h + cs * k.h, xs.u + cs * k.u, xs.v + cs * k.v};
      }
    } else if (live && t >= c0 && t < c1 && owner) {
      const size_t q = static_cast<size_t>(t) * nx;
      oh[q] = xs.h + sdt * (a.h + k.h);
      ou[q] = xs.u + sdt * (a.u + k.u);
      ov[q] = xs.v + sdt * (a.v + k.v);
      if (!INNER && t == ny - 1) ov[static_cast<size_t>(ny) * nx] = M.x[1].v;  // the northern wall row of v
    }
  }
}

// The interior rows of the march with the windows as rings instead of shifted arrays: row rho of the state, of the
// RK4 sums and of f sits in slot rho mod 4, a stage input row in slot rho mod 2, the prefetched state row in slot
// rho mod 2.  With PH = r mod 4 known at compile time every slot index is a constant, so four iterations unrolled
// move no registers (the shifts took 12% of the FP32 interior loop's instructions).  The arithmetic is
// march_iteration<true>'s.
template <class T> struct Ring {
  Row<T> x[4], n[2], acc[4], y[3][2];
  T f[4];
};
template <int PH, int S, class T>
__device__ inline void ring_stage(Ring<T> &R, StageCarry<T> &w, const Stencil<T> &S_, Row<T> &p, int r, int ny, int nx, int c0, int c1, bool owner,
                                  T hdt, T fdt, T sdt, T *__restrict__ oh, T *__restrict__ ou, T *__restrict__ ov) {
LMW_PLANE_CONTRACT
  constexpr int XS = (PH - 1 - S) & 3;  // slot of row t = r - 1 - S
  const int t = r - 1 - S;
  Row<T> k;
  const T f1 = R.f[(PH - S) & 3];
  if constexpr (S == 0) march_row<true>(S_, w, R.x[(PH - 2) & 3], R.x[XS], p, t, ny, f1, k);
  else {
    march_row<true>(S_, w, R.y[S - 1][(PH - S) & 1], R.y[S - 1][(PH - 1 - S) & 1], p, t, ny, f1, k);
    R.y[S - 1][(PH - S) & 1] = p;  // row t + 1 replaces row t - 1
  }
  const Row<T> &xs = R.x[XS];
  Row<T> &a = R.acc[XS];
  const T two = static_cast<T>(2);
  if constexpr (S < 3) {
    const T cs = S == 2 ? fdt : hdt;
    if constexpr (S == 0) a = k;
    else a = Row<T>{a.h + two * k.h, a.u + two * k.u, a.v + two * k.v};
    p = Row<T>{xs.h + cs * k.h, xs.u + cs * k.u, xs.v + cs * k.v};
  } else if (t >= c0 && t < c1 && owner) {
    const size_t q = static_cast<size_t>(t) * nx;
    oh[q] = xs.h + sdt * (a.h + k.h);
    ou[q] = xs.u + sdt * (a.u + k.u);
    ov[q] = xs.v + sdt * (a.v + k.v);
  }
}
template <int PH, class T>
__device__ inline void ring_iteration(Ring<T> &R, StageCarry<T> (&w)[4], const Stencil<T> &S, const T *__restrict__ fcor, int r, int ny, int nx,
                                      int c0, int c1, bool owner, T hdt, T fdt, T sdt, const T *__restrict__ gh, const T *__restrict__ gu,
                                      const T *__restrict__ gv, T *__restrict__ oh, T *__restrict__ ou, T *__restrict__ ov) {
  const Row<T> fresh = R.n[PH & 1];
  {
    const size_t q = static_cast<size_t>(r + 1) * nx;  // r + 1 < ny in the interior loop
    R.n[(PH + 1) & 1] = Row<T>{gh[q], gu[q], gv[q]};
  }
  Row<T> p = fresh;
  ring_stage<PH, 0>(R, w[0], S, p, r, ny, nx, c0, c1, owner, hdt, fdt, sdt, oh, ou, ov);
  ring_stage<PH, 1>(R, w[1], S, p, r, ny, nx, c0, c1, owner, hdt, fdt, sdt, oh, ou, ov);
  ring_stage<PH, 2>(R, w[2], S, p, r, ny, nx, c0, c1, owner, hdt, fdt, sdt, oh, ou, ov);
  ring_stage<PH, 3>(R, w[3], S, p, r, ny, nx, c0, c1, owner, hdt, fdt, sdt, oh, ou, ov);
  R.x[PH & 3] = fresh;               // state row r replaces row r - 4
  R.f[(PH + 1) & 3] = fcor[r + 1];  // f at corner row r + 1 replaces r - 3
}
// The windows between iterations, from march_iteration's shifted arrays to the rings (before iteration r, r = 0
// mod 4) and back (before iteration r, r = 0 mod 4, after the last ring iteration).
template <class T> __device__ inline void to_ring(const March<T> &M, Ring<T> &R) {
#pragma unroll
  for (int i = 0; i < 4; i++) R.x[i] = M.x[i + 1];
  R.n[0] = M.next;
#pragma unroll
  for (int i = 0; i < 3; i++) R.acc[i] = M.acc[i + 1];
  R.f[0] = M.fnext;
#pragma unroll
  for (int i = 1; i < 4; i++) R.f[i] = M.f[i];
#pragma unroll
  for (int s = 1; s < 4; s++) {
    R.y[s - 1][s & 1] = M.ym[s - 1];
    R.y[s - 1][(s + 1) & 1] = M.yc[s - 1];
  }
}
template <class T> __device__ inline void from_ring(const Ring<T> &R, March<T> &M) {
#pragma unroll
  for (int i = 0; i < 4; i++) M.x[i + 1] = R.x[i];
  M.next = R.n[0];
  M.acc[0] = R.acc[3];
#pragma unroll
  for (int i = 0; i < 3; i++) M.acc[i + 1] = R.acc[i];
#pragma unroll
  for (int i = 1; i < 4; i++) M.f[i] = R.f[i];
  M.fnext = R.f[0];
#pragma unroll
  for (int s = 1; s < 4; s++) {
    M.ym[s - 1] = R.y[s - 1][s & 1];
    M.yc[s - 1] = R.y[s - 1][(s + 1) & 1];
  }
}

// One whole RK4 step (plane.cpp:143-158) by row march: a wavefront holds a strip of 64 columns (56 of them its own,
// 4 on each side recomputed for the neighbours' stencils) and walks north through rows c0 - 4 .. c1 + 3 of one
// member, loading each state row once.  Stage s works on row r - s while row r is loaded, so every stage's input at
// the row above is the previous stage's fresh output; all windows are registers, x neighbours come from the adjacent
// lanes, and nothing is shared between wavefronts.  The rows away from the walls run a loop without wall tests.  Per
// cell and stage the arithmetic is that of k_plane_stage, in the same order: the step is bit-identical to four stage
// launches.
// GENERAL: linear or viscous physics (runtime flags); else nonlinear and inviscid.  WAVES: wavefronts per SIMD the
// register allocation must allow (the few values that do not fit are spilled outside the interior loop).
template <class T, bool GENERAL, int WAVES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_plane_march(PlaneGeo P, double dt, int rows,
                                                                                              const T *__restrict__ fcor,
                                                                                              const T *__restrict__ in, T *__restrict__ out) {
  constexpr int HALO = 4, COLS = 64 - 2 * HALO;
  const int nx = P.nx, ny = P.ny, lane = threadIdx.x, b = blockIdx.z;
  const int first = blockIdx.x * COLS - HALO + lane;  // column before wrapping
  const int col = (first % nx + nx) % nx;
  const bool owner = lane >= HALO && lane < HALO + COLS && first < nx;
  const int c0 = blockIdx.y * rows, c1 = min(c0 + rows, ny), r0 = max(c0 - HALO, 0), r1 = c1 + HALO - 1;
  const size_t nh = static_cast<size_t>(nx) * ny, per = 2 * nh + static_cast<size_t>(nx) * (ny + 1);
  const T *gh = in + b * per + col, *gu = gh + nh, *gv = gu + nh;
  T *oh = out + b * per + col, *ou = oh + nh, *ov = ou + nh;
  Stencil<T> S(P);
  const T hdt = static_cast<T>(0.5 * dt), fdt = static_cast<T>(dt), sdt = static_cast<T>(dt / 6);
  if (!GENERAL) {  // nonlinear and inviscid: the physics branches fold away
    S.lin = false;
    S.viscous = false;
  }
  March<T> M = {};
  {
    const size_t q = static_cast<size_t>(r0) * nx;  // r0 < ny
    M.next = Row<T>{gh[q], gu[q], gv[q]};
    M.fnext = fcor[r0];
  }
  // interior rows: every stage's row in 2..ny-2 and row r + 1 exists (FP32: from a multiple of 4, in groups of four)
  const int lo = sizeof(T) == 4 && !GENERAL ? (max(r0, 6) + 3) & ~3 : max(r0, 6), hi = min(r1 - 1, ny - 2);
  int r = r0;
  for (; r <= r1 && r < lo; r++) march_iteration<false>(M, S, fcor, r, r1, ny, nx, c0, c1, owner, hdt, fdt, sdt, gh, gu, gv, oh, ou, ov);
  // (FP32 only: in FP64 the unrolled rings need more than the 168 VGPRs of three wavefronts per SIMD, and two
  // wavefronts per SIMD without the register moves ran no faster than three with them)
  if constexpr (sizeof(T) == 4 && !GENERAL) {
    if (r + 3 <= hi) {
      Ring<T> R;
      to_ring(M, R);
      for (; r + 3 <= hi; r += 4) {
        ring_iteration<0>(R, M.w, S, fcor, r, ny, nx, c0, c1, owner, hdt, fdt, sdt, gh, gu, gv, oh, ou, ov);
        ring_iteration<1>(R, M.w, S, fcor, r + 1, ny, nx, c0, c1, owner, hdt, fdt, sdt, gh, gu, gv, oh, ou, ov);
        ring_iteration<2>(R, M.w, S, fcor, r + 2, ny, nx, c0, c1, owner, hdt, fdt, sdt, gh, gu, gv, oh, ou, ov);
        ring_iteration<3>(R, M.w, S, fcor, r + 3, ny, nx, c0, c1, owner, hdt, fdt, sdt, gh, gu, gv, oh, ou, ov);
      }
      from_ring(R, M);
    }
  }
  for (; r <= hi; r++) march_iteration<true>(M, S, fcor, r, r1, ny, nx, c0, c1, owner, hdt, fdt, sdt, gh, gu, gv, oh, ou, ov);
  for (; r <= r1; r++) march_iteration<false>(M, S, fcor, r, r1, ny, nx, c0, c1, owner, hdt, fdt, sdt, gh, gu, gv, oh, ou, ov);
}

// f at the corner rows 0..ny, as the stencil computes it
template <class T> __global__ void k_coriolis(PlaneGeo P, T *__restrict__ f) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j <= P.ny) f[j] = coriolis<T>(P, j);
}

// Tendency of one member into dh, du, dv (dv has ny+1 rows; rows 0 and ny are zero).
template <class T>
__global__ __launch_bounds__(256) void k_plane_tendency(PlaneGeo P, const T *__restrict__ in, T *__restrict__ dh, T *__restrict__ du, T *__restrict__ dv) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= P.nx || j >= P.ny) return;
  const size_t nh = static_cast<size_t>(P.nx) * P.ny;
  T a, b, c;
  cell_tendency<T>(Stencil<T>(P), P, in, in + nh, in + 2 * nh, i, j, a, b, c);
  const size_t q = static_cast<size_t>(j) * P.nx + i;
  dh[q] = a;
  du[q] = b;
  dv[q] = c;
  if (j == P.ny - 1) dv[static_cast<size_t>(P.ny) * P.nx + i] = 0;
}

struct Kahan {
  double s = 0, c = 0;
  __device__ void add(double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  __device__ double value() const { return s + c; }
};

// Diagnostics of each member (plane.cpp:191-243, and the finiteness sum of plane.cpp:176-181), in two passes.  Pass
// 1, one block per (row j, member b): rows[b][j][q] for q = 0 sum h, 1 sum of energy density, 2 max speed, 3 max h,
// 4 sum pv'^2 over the interior corners of row j, 5 sum h + u + v (v's northern wall row with the last row); the
// threads stride over the columns with compensated sums, combined in thread order.
template <class T>
__global__ __launch_bounds__(256) void k_plane_rows(PlaneGeo P, const T *__restrict__ x, double *__restrict__ rows) {
  const int j = blockIdx.x, b = blockIdx.y, t = threadIdx.x, nx = P.nx, ny = P.ny;
  const size_t nh = static_cast<size_t>(nx) * ny, nv = static_cast<size_t>(nx) * (ny + 1), per = 2 * nh + nv;
  const T *h = x + b * per, *u = h + nh, *v = u + nh;
  __shared__ double part[6][256];
  Kahan rh, re, rpv, rf;
  double smax = 0, hmax = 0;
  for (int i = t; i < nx; i += blockDim.x) {
    const size_t c = static_cast<size_t>(j) * nx + i;
    const int ip = i == nx - 1 ? 0 : i + 1, im = i == 0 ? nx - 1 : i - 1;
    const double hh = h[c], uc = 0.5 * (static_cast<double>(u[c]) + u[static_cast<size_t>(j) * nx + ip]);
    const double vc = 0.5 * (static_cast<double>(v[c]) + v[c + nx]);
    rh.add(hh);
    re.add(0.5 * hh * (uc * uc + vc * vc) + 0.5 * P.g * hh * hh);
    smax = fmax(smax, hypot(uc, vc));
    hmax = fmax(hmax, hh);
    rf.add(hh + static_cast<double>(u[c]) + static_cast<double>(v[c]));
    if (j >= 1) {  // linear PV anomaly at the interior corner (i, j) (plane.cpp:228-243)
      const double hav = 0.25 * (static_cast<double>(h[c - nx + (im - i)]) + h[c - nx] + h[c + (im - i)] + h[c]);
      const double zeta = (static_cast<double>(v[c]) - v[c + (im - i)]) / P.dx - (static_cast<double>(u[c]) - u[c - nx]) / P.dy;
      const double q = zeta - P.f0 * (hav - P.depth) / P.depth;
      rpv.add(q * q);
    }
    if (j == ny - 1) rf.add(static_cast<double>(v[static_cast<size_t>(ny) * nx + i]));
  }
  part[0][t] = rh.value();
  part[1][t] = re.value();
  part[2][t] = smax;
  part[3][t] = hmax;
  part[4][t] = rpv.value();
  part[5][t] = rf.value();
  __syncthreads();
  if (t < 6) {
    double r;
    if (t == 2 || t == 3) {
      r = 0;
      for (int q = 0; q < blockDim.x; q++) r = fmax(r, part[t][q]);
    } else {
      Kahan s;
      for (int q = 0; q < blockDim.x; q++) s.add(part[t][q]);
      r = s.value();
    }
    rows[(static_cast<size_t>(b) * ny + j) * 6 + t] = r;
  }
}
// Pass 2, one block per member: the row values combined in row order into out[b][0..5], staged through LDS 256
// rows at a time so that the six sequential sums read LDS.
__global__ __launch_bounds__(256) void k_plane_rows_combine(int ny, const double *__restrict__ rows, double *__restrict__ out) {
  const int b = blockIdx.x, t = threadIdx.x;
  __shared__ double chunk[256 * 6];
  const double *r = rows + static_cast<size_t>(b) * ny * 6;
  double m = 0;
  Kahan s;
  for (int j0 = 0; j0 < ny; j0 += 256) {
    const int n = min(256, ny - j0);
    for (int e = t; e < n * 6; e += blockDim.x) chunk[e] = r[static_cast<size_t>(j0) * 6 + e];
    __syncthreads();
    if (t < 6)
      for (int j = 0; j < n; j++) {
        if (t == 2 || t == 3) m = fmax(m, chunk[j * 6 + t]);
        else s.add(chunk[j * 6 + t]);
      }
    __syncthreads();
  }
  if (t < 6) out[static_cast<size_t>(b) * 8 + t] = (t == 2 || t == 3) ? m : s.value();
}

// The zonal harmonic m of rows j0 .. j0 + gridDim.x - 1 of one field of member b (plane.cpp:264-268):
// out[r] = (1/nx) sum_i row[i] exp(-2 pi i m (i + shift2 / 2) / nx).  The phase is reduced exactly in integers,
// m (2 i + shift2) mod 2 nx, before sincospi, so no angle grows with m or i.  One block per row; each thread's
// compensated sums, combined in thread order.
template <class T>
__global__ __launch_bounds__(256) void k_plane_harmonic(PlaneGeo P, const T *__restrict__ x, int b, int field, int m, int shift2, int j0,
                                                        double *__restrict__ out) {
  const int j = j0 + blockIdx.x, t = threadIdx.x, nx = P.nx, ny = P.ny;
  const size_t nh = static_cast<size_t>(nx) * ny, nv = static_cast<size_t>(nx) * (ny + 1), per = 2 * nh + nv;
  const T *row = x + b * per + (field == 0 ? 0 : field == 1 ? nh : 2 * nh) + static_cast<size_t>(j) * nx;
  __shared__ double part[2][256];
  Kahan re, im;
  const int64_t period = 2LL * nx;
  for (int i = t; i < nx; i += blockDim.x) {
    const int64_t k = (static_cast<int64_t>(m) * (2 * i + shift2)) % period;
    double sn, cs;
    sincospi(static_cast<double>(k) / nx, &sn, &cs);
    const double f = row[i];
    re.add(f * cs);
    im.add(-f * sn);
  }
  part[0][t] = re.value();
  part[1][t] = im.value();
  __syncthreads();
  if (t < 2) {
    Kahan s;
    for (int q = 0; q < blockDim.x; q++) s.add(part[t][q]);
    out[2 * static_cast<size_t>(blockIdx.x) + t] = s.value() / nx;
  }
}

template <class T> __global__ void k_convert(const double *__restrict__ src, T *__restrict__ dst, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = static_cast<T>(src[i]);
}
template <class T> __global__ void k_widen(const T *__restrict__ src, double *__restrict__ dst, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = static_cast<double>(src[i]);
}

// Plane.fields (cpp/main.cpp:410-431) on the device: h = H + e0 at the cell centres, u = e1 at the west faces, v = e2
// at the interior south faces (rows 1 .. ny-1; the wall rows 0 and ny are 0), written in FP64 into the staging
// state io = [h][u][v].  xy = [nx] centre x, [nx] face x, [ny] centre y, [ny + 1] face y.  bad[f] is set where
// field f is not finite.
template <int D>
__global__ __launch_bounds__(256) void k_plane_expr(ExprSet X, const double *__restrict__ xy, int nx, int ny, double depth,
                                                    double *__restrict__ io, int *__restrict__ bad) {
  const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (i >= nx) return;
  const double *xc = xy, *xu = xy + nx, *yc = xy + 2 * nx, *yv = xy + 2 * nx + ny;
  const size_t nh = static_cast<size_t>(nx) * ny, at = static_cast<size_t>(j) * nx + i;
  if (j < ny) {
    const double h = depth + expr_eval<D>(X.code + X.start[0], X.len[0], xc[i], yc[j]);
    const double u = expr_eval<D>(X.code + X.start[1], X.len[1], xu[i], yc[j]);
    if (!isfinite(h)) bad[0] = 1;
    if (!isfinite(u)) bad[1] = 1;
    io[at] = h;
    io[nh + at] = u;
  }
  double v = 0;
  if (j >= 1 && j < ny) v = expr_eval<D>(X.code + X.start[2], X.len[2], xc[i], yv[j]);
  if (!isfinite(v)) bad[2] = 1;
  io[2 * nh + at] = v;
}

// The seeded noise of perturb (plane.cpp:34-57) on the device.  splitmix64 advances its state by a constant, so the
// k-th draw of the stream is mix(st0 + k gamma) and every cell's twelve draws (12 c + 1 .. 12 c + 12) are computed
// independently; the twelve terms are summed in draw order, as the sequential stream sums them, so the noise is bit
// for bit the reference's.  These kernels keep contraction off in both precisions: the perturbation is the same
// double-precision field before the state's rounding to T.
__device__ inline uint64_t splitmix_at(uint64_t st0, uint64_t k) {
  uint64_t z = st0 + k * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__global__ void k_plane_noise(uint64_t st0, size_t cells, double *__restrict__ n) {
  #pragma clang fp contract(off)
  const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  double s = 0;
  for (int k = 0; k < 12; k++) s += static_cast<double>(splitmix_at(st0, 12 * c + k + 1) >> 11) * (1.0 / 9007199254740992.0);
  n[c] = s - 6.0;
}
// One 1-2-1 pass along x (periodic; dir 0) or along y (the edge rows repeated; dir 1) of an ny x nx field.
__global__ void k_plane_smooth(int nx, int ny, int dir, const double *__restrict__ in, double *__restrict__ out) {
  #pragma clang fp contract(off)
  const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= static_cast<size_t>(nx) * ny) return;
  const int j = static_cast<int>(c / nx), i = static_cast<int>(c - static_cast<size_t>(j) * nx);
  double a, z;
  if (dir == 0) {
    a = in[c - i + (i == 0 ? nx - 1 : i - 1)];
    z = in[c - i + (i == nx - 1 ? 0 : i + 1)];
  } else {
    a = in[j == 0 ? c : c - nx];
    z = in[j == ny - 1 ? c : c + nx];
  }
  out[c] = 0.25 * a + 0.5 * in[c] + 0.25 * z;
}
// h += scale (n - mean), in double, rounded to T once
template <class T> __global__ void k_plane_add_noise(T *__restrict__ h, const double *__restrict__ n, double mean, double scale, size_t cells) {
  #pragma clang fp contract(off)
  const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < cells) h[c] = static_cast<T>(static_cast<double>(h[c]) + (n[c] - mean) * scale);
}

}  // namespace

// A batch of planes resident on the GPU; T is the state and arithmetic type.
struct PlaneCoreBase {
  virtual ~PlaneCoreBase() = default;
  lmw_plane_params prm;
  int B = 1;
  double t = 0;
  hipStream_t stream = nullptr;
  size_t nh() const { return static_cast<size_t>(prm.nx) * prm.ny; }
  size_t nv() const { return static_cast<size_t>(prm.nx) * (prm.ny + 1); }
  size_t per() const { return 2 * nh() + nv(); }
  PlaneGeo geo() const {
    PlaneGeo P;
    P.nx = prm.nx;
    P.ny = prm.ny;
    P.B = B;
    P.dx = prm.lx / prm.nx;
    P.dy = prm.ly / prm.ny;
    P.g = prm.gravity;
    P.depth = prm.depth;
    P.f0 = prm.f0;
    P.beta = prm.beta;
    P.ly = prm.ly;
    P.visc = prm.viscosity;
    P.drag = prm.drag;
    P.linear = prm.linear;
    return P;
  }
  virtual void set_field(int b, int field, const double *g) = 0;
  virtual void get_field(int b, int field, double *out) = 0;
  virtual void steps(double dt, int64_t count, double *kernel_ms = nullptr) = 0;
  virtual void reduce(double *out8B) = 0;  // B x 8
  virtual void tendency(int b, double *dh, double *du, double *dv) = 0;
  // the zonal harmonic m of rows j0 .. j0 + rows - 1 of one field of member b, (re, im) per row into out[2 * rows]
  virtual void harmonic(int b, int field, int m, int shift2, int j0, int rows, double *out) = 0;
  virtual void set_tile(int tile) = 0;
  virtual void reset() = 0;
  // perturb (plane.cpp:34-57) of member b's h
  virtual void perturb(int b, double amplitude, uint64_t seed) = 0;
  // Plane.fields (cpp/main.cpp:410-431) of member b from expressions in x, y and lx, ly, H, g, f0, beta
  virtual void fields_expr(int b, const std::string &h, const std::string &u, const std::string &v) = 0;
  void sync() { LMW_HIP(hipStreamSynchronize(stream)); }
};

// The two precisions are built in separate translation units (plane.hip FP64, plane_f32.hip FP32), each with its
// own LMW_PLANE_CONTRACT, and created through these factories.
std::unique_ptr<PlaneCoreBase> make_plane_core_f64(const lmw_plane_params &p, int members);
std::unique_ptr<PlaneCoreBase> make_plane_core_f32(const lmw_plane_params &p, int members);

namespace {

template <class T> struct PlaneCore : PlaneCoreBase {
  DevArray<T> x0, xa, xb, acc;
  DevArray<double> red, io;
  DevArray<T> tmp;
  PlaneCore(const lmw_plane_params &p, int members) {
    prm = p;
    B = members;
    LMW_HIP(hipStreamCreate(&stream));
    const size_t n = static_cast<size_t>(B) * per();
    x0.alloc(n);
    xa.alloc(n);
    xb.alloc(n);
    acc.alloc(n);
    red.alloc(static_cast<size_t>(B) * 8);
    io.alloc(per());
    tmp.alloc(3 * per());
    reset();
  }
  ~PlaneCore() override {
    if (stream) (void)hipStreamDestroy(stream);
  }
  size_t field_off(int b, int field) const { return static_cast<size_t>(b) * per() + (field == 0 ? 0 : field == 1 ? nh() : 2 * nh()); }
  size_t field_len(int field) const { return field == 2 ? nv() : nh(); }
  DevArray<double> noise;  // [2][ny][nx]: the noise field and the smoothing passes' other buffer
  void perturb(int b, double amplitude, uint64_t seed) override {
    const size_t cells = nh();
    if (noise.n < 2 * cells) noise.alloc(2 * cells);
    const unsigned blocks = static_cast<unsigned>((cells + 255) / 256);
    double *n = noise.p, *t = noise.p + cells;
    hipLaunchKernelGGL(k_plane_noise, dim3(blocks), dim3(256), 0, stream, seed * 0x9E3779B97F4A7C15ULL + 777, cells, n);
    LMW_LAUNCH_CHECK();
    for (int pass = 0; pass < 3; pass++) {
      hipLaunchKernelGGL(k_plane_smooth, dim3(blocks), dim3(256), 0, stream, prm.nx, prm.ny, 0, n, t);
      LMW_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_plane_smooth, dim3(blocks), dim3(256), 0, stream, prm.nx, prm.ny, 1, t, n);
      LMW_LAUNCH_CHECK();
    }
    // the mean and the rms in the reference's sequential order (host), so the scale is its scale bit for bit
    std::vector<double> hn(cells);
    noise.download(hn.data(), cells, 0, stream);
    sync();
    double mean = 0, ms = 0;
    for (double x : hn) mean += x;
    mean /= static_cast<double>(cells);
    for (double x : hn) {
      const double d = x - mean;
      ms += d * d;
    }
    const double rms = std::sqrt(ms / static_cast<double>(cells));
    const double scale = rms > 0 ? amplitude / rms : 0;
    hipLaunchKernelGGL(k_plane_add_noise<T>, dim3(blocks), dim3(256), 0, stream, x0.p + field_off(b, 0), n, mean, scale, cells);
    LMW_LAUNCH_CHECK();
    sync();
  }
  ExprDevice expr;
  DevArray<double> xy;
  void fields_expr(int b, const std::string &h, const std::string &u, const std::string &v) override {
    const std::vector<std::pair<std::string, double>> par = {{"lx", prm.lx}, {"ly", prm.ly}, {"H", prm.depth},
                                                             {"g", prm.gravity}, {"f0", prm.f0}, {"beta", prm.beta}};
    const ExprProgram ph = compile_expr(h, "x", "y", par), pu = compile_expr(u, "x", "y", par), pv = compile_expr(v, "x", "y", par);
    const ExprProgram *progs[3] = {&ph, &pu, &pv};
    const ExprSet X = expr.load(progs, 3, stream);
    // the coordinates as the reference forms them on the host (plane.hpp:26-31, main.cpp:418-422)
    const int nx = prm.nx, ny = prm.ny;
    const double dx = prm.lx / nx, dy = prm.ly / ny;
    std::vector<double> c(2 * static_cast<size_t>(nx) + 2 * ny + 1);
    for (int i = 0; i < nx; i++) {
      c[i] = (i + 0.5) * dx;
      c[nx + i] = i * dx;
    }
    for (int j = 0; j < ny; j++) c[2 * nx + j] = (j + 0.5) * dy;
    for (int j = 0; j <= ny; j++) c[2 * nx + ny + j] = j * dy;
    if (xy.n < c.size()) xy.alloc(c.size());
    xy.upload(c.data(), c.size(), 0, stream);
    const dim3 grd((nx + 255) / 256, ny + 1);
    if (expr.depth <= kExprSmallDepth)
      hipLaunchKernelGGL(k_plane_expr<kExprSmallDepth>, grd, dim3(256), 0, stream, X, xy.p, nx, ny, prm.depth, io.p, expr.bad.p);
    else
      hipLaunchKernelGGL(k_plane_expr<kExprMaxDepth>, grd, dim3(256), 0, stream, X, xy.p, nx, ny, prm.depth, io.p, expr.bad.p);
    LMW_LAUNCH_CHECK();
    int bad[3];
    expr.bad.download(bad, 3, 0, stream);
    sync();
    if (bad[0]) runtime("Plane.fields: the height expression is not finite");
    if (bad[1]) runtime("Plane.fields: the u expression is not finite");
    if (bad[2]) runtime("Plane.fields: the v expression is not finite");
    const size_t n = per();  // h, u, v of member b are contiguous, in io's order
    hipLaunchKernelGGL(k_convert<T>, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, io.p, x0.p + field_off(b, 0), n);
    LMW_LAUNCH_CHECK();
    sync();
  }
  void reset() override {
    // h = depth, u = v = 0 (plane.cpp:27-32)
    std::vector<double> hh(nh(), prm.depth), zero(nv(), 0.0);
    for (int b = 0; b < B; b++) {
      set_field(b, 0, hh.data());
      set_field(b, 1, zero.data());
      set_field(b, 2, zero.data());
    }
    t = 0;
  }
  void set_field(int b, int field, const double *g) override {
    const size_t n = field_len(field);
    io.upload(g, n, 0, stream);
    const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
    hipLaunchKernelGGL(k_convert<T>, dim3(blocks), dim3(256), 0, stream, io.p, x0.p + field_off(b, field), n);
    LMW_LAUNCH_CHECK();
    sync();
  }
  void get_field(int b, int field, double *out) override {
    const size_t n = field_len(field);
    const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
    hipLaunchKernelGGL(k_widen<T>, dim3(blocks), dim3(256), 0, stream, x0.p + field_off(b, field), io.p, n);
    LMW_LAUNCH_CHECK();
    io.download(out, n, 0, stream);
    sync();
  }
  void launch_stage(int stage, double dt, const T *in, T *out) {
    dim3 grd((prm.nx + 63) / 64, (prm.ny + 3) / 4, B), blk(64, 4);
    hipLaunchKernelGGL(k_plane_stage<T>, grd, blk, 0, stream, geo(), stage, dt, in, x0.p, out, acc.p);
    LMW_LAUNCH_CHECK();
  }
  // whole RK4 steps, one row-march launch each (k_plane_march), the state ping-ponging between x0 and xa; f by row
  // in fcor.  tile 0 = rows per wavefront from the occupancy, 1..4 = 32, 64, 128, 256 rows, 5 = four stage launches
  int tile = 0;
  DevArray<T> fcor;
  static constexpr int kMarchWaves = sizeof(T) == 8 ? 3 : 4;  // 168 / 128 VGPRs: no spills in the interior loops
  int auto_rows = 0;
  int march_rows() {
    if (tile >= 1 && tile <= 4) return 32 << (tile - 1);
    if (auto_rows) return auto_rows;
    // the largest chunk that still gives every resident wavefront slot a wavefront (longer chunks recompute fewer
    // halo rows; measured at 1024 x 512 x 64: 128 rows in FP64, 64 in FP32)
    int dev = 0, cus = 1, per_cu = 1;
    LMW_HIP(hipGetDevice(&dev));
    LMW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    LMW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_plane_march<T, false, kMarchWaves>, 64, 0));
    const long slots = static_cast<long>(cus) * std::max(per_cu, 1), strips = (prm.nx + 55) / 56;
    int rows = 256;
    while (rows > 32 && strips * ((prm.ny + rows - 1) / rows) * static_cast<long>(B) < slots) rows /= 2;
    return auto_rows = rows;
  }
  void launch_march(double dt, int rows) {
    dim3 grd((prm.nx + 55) / 56, (prm.ny + rows - 1) / rows, B), blk(64);
    if (prm.linear || prm.viscosity != 0) hipLaunchKernelGGL((k_plane_march<T, true, 1>), grd, blk, 0, stream, geo(), dt, rows, fcor.p, x0.p, xa.p);
    else hipLaunchKernelGGL((k_plane_march<T, false, kMarchWaves>), grd, blk, 0, stream, geo(), dt, rows, fcor.p, x0.p, xa.p);
    LMW_LAUNCH_CHECK();
    std::swap(x0.p, xa.p);
  }
  void set_tile(int t) override {
    if (t < 0 || t > 5) invalid("plane step kernel must be 0 .. 5");
    tile = t;
  }
  // kernel_ms: if given, receives the time of the step launches alone (HIP events on the plane's stream), in ms
  void steps(double dt, int64_t count, double *kernel_ms) override {
    if (kernel_ms) *kernel_ms = 0;
    if (count <= 0) return;
    int rows = 0;
    if (tile != 5) {
      rows = march_rows();
      if (fcor.n < static_cast<size_t>(prm.ny) + 1) fcor.alloc(static_cast<size_t>(prm.ny) + 1);
      hipLaunchKernelGGL(k_coriolis<T>, dim3(prm.ny / 256 + 1), dim3(256), 0, stream, geo(), fcor.p);
      LMW_LAUNCH_CHECK();
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (kernel_ms) {
      for (auto &e : ev) LMW_HIP(hipEventCreate(&e));
      LMW_HIP(hipEventRecord(ev[0], stream));
    }
    if (tile == 5) steps_staged(dt, count);
    else
      for (int64_t s = 0; s < count; s++) launch_march(dt, rows);
    if (kernel_ms) {
      LMW_HIP(hipEventRecord(ev[1], stream));
      LMW_HIP(hipEventSynchronize(ev[1]));
      float ms = 0;
      LMW_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
      *kernel_ms = ms;
      for (auto &e : ev) (void)hipEventDestroy(e);
    }
  }
  // the same steps as four stage launches through x0, acc and the stage buffers
  void steps_staged(double dt, int64_t count) {
    for (int64_t s = 0; s < count; s++) {
      launch_stage(1, dt, x0.p, xa.p);
      launch_stage(2, dt, xa.p, xb.p);
      launch_stage(3, dt, xb.p, xa.p);
      launch_stage(4, dt, xa.p, nullptr);
    }
  }
  DevArray<double> rowred;
  void reduce(double *out) override {
    if (rowred.n < static_cast<size_t>(B) * prm.ny * 6) rowred.alloc(static_cast<size_t>(B) * prm.ny * 6);
    hipLaunchKernelGGL(k_plane_rows<T>, dim3(prm.ny, B), dim3(256), 0, stream, geo(), x0.p, rowred.p);
    LMW_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_plane_rows_combine, dim3(B), dim3(256), 0, stream, prm.ny, rowred.p, red.p);
    LMW_LAUNCH_CHECK();
    red.download(out, static_cast<size_t>(B) * 8, 0, stream);
    sync();
  }
  DevArray<double> harm;
  void harmonic(int b, int field, int m, int shift2, int j0, int rows, double *out) override {
    if (rows < 1) return;
    if (harm.n < 2 * static_cast<size_t>(rows)) harm.alloc(2 * static_cast<size_t>(rows));
    hipLaunchKernelGGL(k_plane_harmonic<T>, dim3(rows), dim3(256), 0, stream, geo(), x0.p, b, field, m, shift2, j0, harm.p);
    LMW_LAUNCH_CHECK();
    harm.download(out, 2 * static_cast<size_t>(rows), 0, stream);
    sync();
  }
  void tendency(int b, double *dh, double *du, double *dv) override {
    dim3 grd((prm.nx + 63) / 64, (prm.ny + 3) / 4, 1), blk(64, 4);
    T *th = tmp.p, *tu = th + per(), *tv = tu + per();
    hipLaunchKernelGGL(k_plane_tendency<T>, grd, blk, 0, stream, geo(), x0.p + static_cast<size_t>(b) * per(), th, tu, tv);
    LMW_LAUNCH_CHECK();
    double *outs[3] = {dh, du, dv};
    T *srcs[3] = {th, tu, tv};
    const size_t lens[3] = {nh(), nh(), nv()};
    for (int f = 0; f < 3; f++) {
      hipLaunchKernelGGL(k_widen<T>, dim3(static_cast<unsigned>((lens[f] + 255) / 256)), dim3(256), 0, stream, srcs[f], io.p, lens[f]);
      LMW_LAUNCH_CHECK();
      io.download(outs[f], lens[f], 0, stream);
      sync();
    }
  }
};

}  // namespace

}  // namespace lmw
