// Arakawa C-grid shallow water on an f/beta-plane channel for gfx950 (reference cpp/plane.cpp), batched over
// members that share the parameters, in FP64 or FP32.
//
// Layout: per member h[ny][nx], u[ny][nx] (west faces), v[ny+1][nx] (south faces; rows 0 and ny are walls, zero),
// members contiguous.  The step path is k_plane_step: one launch per RK4 step, each block loading a tile plus a
// 4-cell halo into LDS, running the four stages there (Bernoulli function at centres and absolute vorticity at
// corners once per stage, then du, dv, dh per cell: plane.cpp:59-129, and the stage update: plane.cpp:143-158), and
// writing its interior back, so HBM sees the state once in and once out per step.  k_plane_stage, one launch per
// stage with a 64 x 4 tile per block and buffers x0 (state), xa, xb (stage inputs, ping-pong) and acc (RK4 sum),
// computes the same arithmetic in the same order and is kept as the reference shape for the fused kernel.
#include <algorithm>
#include <cmath>
#include <memory>

#include "common.h"

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
// An LDS tile of EX columns and rows jb .. jb + EY - 1 (v: jb .. jb + EY): neighbours outside the tile are clamped
// to its edge, which only affects cells whose results the caller discards.
template <class T> struct TileAccess {
  const T *h, *u, *v;
  int ex, ey, jb;
  __device__ int row(int j, int last) const { return min(max(j - jb, 0), last); }
  __device__ T H(int i, int j) const { return h[row(j, ey - 1) * ex + i]; }
  __device__ T U(int i, int j) const { return u[row(j, ey - 1) * ex + i]; }
  __device__ T V(int i, int j) const { return v[row(j, ey) * ex + i]; }
  __device__ int xm(int i) const { return i > 0 ? i - 1 : 0; }
  __device__ int xp(int i) const { return i < ex - 1 ? i + 1 : ex - 1; }
};

// B at the centres and eta at the corners of an LDS tile (rows jb .. jb + EY - 1, eta rows to jb + EY)
template <class T> struct TileBE {
  const T *b, *eta;
  int ex, ey, jb;
  __device__ T B(int i, int j) const { return b[min(max(j - jb, 0), ey - 1) * ex + i]; }
  __device__ T Eta(int i, int j) const { return eta[min(max(j - jb, 0), ey) * ex + i]; }
};

// Bernoulli function B = g h + K at centre (ii, jj) (plane.cpp:72-77)
template <class T, class A> __device__ inline T bernoulli(const PlaneGeo &P, const A &F, int ii, int jj) {
#pragma clang fp contract(off)
  const T half = static_cast<T>(0.5), g = static_cast<T>(P.g);
  const int i1 = F.xp(ii);
  T k = 0;
  if (P.linear == 0)
    k = half * (half * (F.U(ii, jj) * F.U(ii, jj) + F.U(i1, jj) * F.U(i1, jj)) + half * (F.V(ii, jj) * F.V(ii, jj) + F.V(ii, jj + 1) * F.V(ii, jj + 1)));
  return g * F.H(ii, jj) + k;
}
// absolute vorticity at corner (ii, jj), jj = 0..ny (plane.cpp:80-92)
template <class T, class A> __device__ inline T abs_vorticity(const PlaneGeo &P, const A &F, int ii, int jj) {
#pragma clang fp contract(off)
  const int ny = P.ny;
  const T dxx = static_cast<T>(P.dx), dyy = static_cast<T>(P.dy);
  const T f = static_cast<T>(P.f0 + P.beta * (jj * P.dy - 0.5 * P.ly));
  T zeta = 0;
  if (P.linear == 0) {
    const int i0 = F.xm(ii);
    const T dvdx = (F.V(ii, jj) - F.V(i0, jj)) / dxx;
    const T uj = jj < ny ? F.U(ii, jj) : F.U(ii, ny - 1), ujm = jj > 0 ? F.U(ii, jj - 1) : F.U(ii, 0);
    zeta = dvdx - (uj - ujm) / dyy;
  }
  return f + zeta;
}
// B and eta computed where they are asked for
template <class T, class A> struct OnTheFly {
  const PlaneGeo &P;
  const A &F;
  __device__ T B(int ii, int jj) const { return bernoulli<T>(P, F, ii, jj); }
  __device__ T Eta(int ii, int jj) const { return abs_vorticity<T>(P, F, ii, jj); }
};

// Tendency of cell (i, j) of one member (plane.cpp:59-129), arithmetic in T; j is the global row (wall logic); E
// supplies the Bernoulli function at centres and the absolute vorticity at corners.
template <class T, class A, class E>
__device__ inline void cell_tendency(const PlaneGeo &P, const A &F, const E &Q, int i, int j, T &dh, T &du, T &dv) {
#pragma clang fp contract(off)
  const int ny = P.ny;
  const T dxx = static_cast<T>(P.dx), dyy = static_cast<T>(P.dy);
  const T half = static_cast<T>(0.5), quarter = static_cast<T>(0.25), two = static_cast<T>(2);
  const int im = F.xm(i), ip = F.xp(i);
  auto H = [&](int ii, int jj) { return F.H(ii, jj); };
  auto U = [&](int ii, int jj) { return F.U(ii, jj); };
  auto V = [&](int ii, int jj) { return F.V(ii, jj); };
  const bool lin = P.linear != 0;
  const T drag = static_cast<T>(P.drag), visc = static_cast<T>(P.visc), depth = static_cast<T>(P.depth);
  // du at the u point (i, j): west face (plane.cpp:94-105)
  {
    const T q = half * (Q.Eta(i, j) + Q.Eta(i, j + 1));
    const T vbar = quarter * (V(im, j) + V(i, j) + V(im, j + 1) + V(i, j + 1));
    T r = q * vbar - (Q.B(i, j) - Q.B(im, j)) / dxx - drag * U(i, j);
    if (P.visc != 0) {
      const T uy1 = j + 1 < ny ? U(i, j + 1) : U(i, j), uy0 = j > 0 ? U(i, j - 1) : U(i, j);
      r += visc * ((U(ip, j) - two * U(i, j) + U(im, j)) / (dxx * dxx) + (uy1 - two * U(i, j) + uy0) / (dyy * dyy));
    }
    du = r;
  }
  // dv at the v point (i, j): south face, rows 1..ny-1 (plane.cpp:106-117)
  if (j >= 1) {
    const T q = half * (Q.Eta(i, j) + Q.Eta(ip, j));
    const T ubar = quarter * (U(i, j - 1) + U(ip, j - 1) + U(i, j) + U(ip, j));
    T r = -q * ubar - (Q.B(i, j) - Q.B(i, j - 1)) / dyy - drag * V(i, j);
    if (P.visc != 0) {
      const T vy1 = j + 1 < ny ? V(i, j + 1) : T(0), vy0 = j - 1 > 0 ? V(i, j - 1) : T(0);
      r += visc * ((V(ip, j) - two * V(i, j) + V(im, j)) / (dxx * dxx) + (vy1 - two * V(i, j) + vy0) / (dyy * dyy));
    }
    dv = r;
  } else {
    dv = 0;
  }
  // continuity in flux form (plane.cpp:119-128)
  {
    const T hw = lin ? depth : half * (H(im, j) + H(i, j));
    const T fe = (lin ? depth : half * (H(i, j) + H(ip, j))) * U(ip, j);
    const T fw = hw * U(i, j);
    T fn = 0, fs = 0;
    if (j + 1 < ny) fn = (lin ? depth : half * (H(i, j) + H(i, j + 1))) * V(i, j + 1);
    if (j > 0) fs = (lin ? depth : half * (H(i, j - 1) + H(i, j))) * V(i, j);
    dh = -((fe - fw) / dxx + (fn - fs) / dyy);
  }
}
template <class T, class A> __device__ inline void cell_tendency(const PlaneGeo &P, const A &F, int i, int j, T &dh, T &du, T &dv) {
  cell_tendency<T>(P, F, OnTheFly<T, A>{P, F}, i, j, dh, du, dv);
}
template <class T>
__device__ inline void cell_tendency(const PlaneGeo &P, const T *__restrict__ h, const T *__restrict__ u, const T *__restrict__ v, int i,
                                     int j, T &dh, T &du, T &dv) {
  cell_tendency<T>(P, GridAccess<T>{h, u, v, P.nx}, i, j, dh, du, dv);
}

// One RK4 stage for all members: tendency of `in`, then the stage update (plane.cpp:143-158).
template <class T>
__global__ __launch_bounds__(256) void k_plane_stage(PlaneGeo P, int stage, double dt, const T *__restrict__ in, T *__restrict__ x0,
                                                     T *__restrict__ out, T *__restrict__ acc) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, b = blockIdx.z;
  if (i >= P.nx || j >= P.ny) return;
  const size_t nh = static_cast<size_t>(P.nx) * P.ny, nv = static_cast<size_t>(P.nx) * (P.ny + 1), per = 2 * nh + nv;
  const T *h = in + b * per, *u = h + nh, *v = u + nh;
  T dh, du, dv;
  cell_tendency<T>(P, h, u, v, i, j, dh, du, dv);
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

// One whole RK4 step (plane.cpp:143-158) of a TX x TY tile per block of NT threads, its four stages in LDS: the
// block loads the tile and a halo of 4 cells (the stencil reaches one cell per stage) from `in`, runs the stages on
// the extended tile (each stage's results are exact one cell further inside than the previous stage's; the clamped
// values at the tile edge only reach cells that are discarded), and writes the interior to `out`.  The stage inputs
// ping-pong between two LDS copies; the RK4 sums of the interior cells stay in registers.  Arithmetic per cell is
// that of k_plane_stage in the same order, so the step is bit-identical to four stage launches, at one read of the
// extended tile and one write of the interior per step.
template <class T, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_plane_step(PlaneGeo P, double dt, const T *__restrict__ in, T *__restrict__ out) {
#pragma clang fp contract(off)
  constexpr int HALO = 4, EX = TX + 2 * HALO, EY = TY + 2 * HALO, NC = EX * EY, NV = EX * (EY + 1), NS = 2 * NC + NV;
  constexpr int PI = TX * TY / NT;  // interior cells per thread
  static_assert(TX * TY % NT == 0, "the interior must divide among the threads");
  extern __shared__ double plane_lds[];
  T *X = reinterpret_cast<T *>(plane_lds), *S0 = X + NS, *S1 = S0 + NS;  // state at the start, two stage inputs
  T *Bc = S1 + NS, *Ec = Bc + NC;  // Bernoulli function at the centres, absolute vorticity at the corners
  const int nx = P.nx, ny = P.ny, b = blockIdx.z, tid = threadIdx.x;
  const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY, ib = i0 - HALO, jb = j0 - HALO;
  const size_t nh = static_cast<size_t>(nx) * ny, nvv = static_cast<size_t>(nx) * (ny + 1), per = 2 * nh + nvv;
  const T *gh = in + b * per, *gu = gh + nh, *gv = gu + nh;
  // load (periodic in x); rows outside the domain are never read by a cell inside it: zeros
  for (int e = tid; e < NV; e += NT) {
    const int ex = e % EX, ey = e / EX, j = jb + ey, i = ((ib + ex) % nx + nx) % nx;
    if (ey < EY) {
      const bool dom = j >= 0 && j < ny;
      const T hv = dom ? gh[static_cast<size_t>(j) * nx + i] : T(0), uv = dom ? gu[static_cast<size_t>(j) * nx + i] : T(0);
      X[e] = S0[e] = hv;
      X[NC + e] = S0[NC + e] = uv;
    }
    const T vv = (j >= 0 && j <= ny) ? gv[static_cast<size_t>(j) * nx + i] : T(0);
    X[2 * NC + e] = S0[2 * NC + e] = S1[2 * NC + e] = vv;  // S1's top v row is never rewritten: it keeps its value
  }
  __syncthreads();
  const T hdt = static_cast<T>(0.5 * dt), fdt = static_cast<T>(dt), sdt = static_cast<T>(dt / 6), two = static_cast<T>(2);
  T acc[PI][3];
  for (int stage = 1; stage <= 4; stage++) {
    const T *cur = (stage & 1) ? S0 : S1;
    T *nxt = (stage & 1) ? S1 : S0;
    const TileAccess<T> F{cur, cur + NC, cur + 2 * NC, EX, EY, jb};
    const T c = stage == 3 ? fdt : hdt;
    // B at every centre and eta at every corner of the tile, once (plane.cpp:72-92)
    for (int e = tid; e < NV; e += NT) {
      const int ex = e % EX, ey = e / EX, j = jb + ey;
      if (ey < EY) Bc[e] = (j >= 0 && j < ny) ? bernoulli<T>(P, F, ex, j) : T(0);
      Ec[e] = (j >= 0 && j <= ny) ? abs_vorticity<T>(P, F, ex, j) : T(0);
    }
    __syncthreads();
    const TileBE<T> Q{Bc, Ec, EX, EY, jb};
    // interior cells: the RK4 sum, then the next stage input or the result
#pragma unroll
    for (int o = 0; o < PI; o++) {
      const int p = tid + NT * o, ix = p % TX, iy = p / TX, e = (iy + HALO) * EX + ix + HALO, j = j0 + iy;
      T k[3] = {T(0), T(0), T(0)};
      if (j < ny) cell_tendency<T>(P, F, Q, ix + HALO, j, k[0], k[1], k[2]);
      for (int f = 0; f < 3; f++) acc[o][f] = stage == 1 ? k[f] : (stage <= 3 ? acc[o][f] + two * k[f] : acc[o][f]);
      if (stage <= 3) {
        for (int f = 0; f < 3; f++) nxt[f * NC + e] = X[f * NC + e] + c * k[f];
      } else if (j < ny && i0 + ix < nx) {
        T *oh = out + b * per, *ou = oh + nh, *ov = ou + nh;
        const size_t q = static_cast<size_t>(j) * nx + i0 + ix;
        oh[q] = X[e] + sdt * (acc[o][0] + k[0]);
        ou[q] = X[NC + e] + sdt * (acc[o][1] + k[1]);
        ov[q] = X[2 * NC + e] + sdt * (acc[o][2] + k[2]);
        if (j == ny - 1) ov[static_cast<size_t>(ny) * nx + i0 + ix] = X[2 * NC + e + EX];  // the northern wall row of v
      }
    }
    // the halo ring: next stage input only
    if (stage <= 3) {
      for (int e = tid; e < NC; e += NT) {
        const int ex = e % EX, ey = e / EX;
        if (ex >= HALO && ex < HALO + TX && ey >= HALO && ey < HALO + TY) continue;
        const int j = jb + ey;
        T k[3] = {T(0), T(0), T(0)};
        if (j >= 0 && j < ny) cell_tendency<T>(P, F, Q, ex, j, k[0], k[1], k[2]);
        for (int f = 0; f < 3; f++) nxt[f * NC + e] = X[f * NC + e] + c * k[f];
      }
    }
    __syncthreads();
  }
}

// Tendency of one member into dh, du, dv (dv has ny+1 rows; rows 0 and ny are zero).
template <class T>
__global__ __launch_bounds__(256) void k_plane_tendency(PlaneGeo P, const T *__restrict__ in, T *__restrict__ dh, T *__restrict__ du, T *__restrict__ dv) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= P.nx || j >= P.ny) return;
  const size_t nh = static_cast<size_t>(P.nx) * P.ny;
  T a, b, c;
  cell_tendency<T>(P, in, in + nh, in + 2 * nh, i, j, a, b, c);
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

// Per member: out[b][0] = sum h, [1] = sum of energy density, [2] = max speed, [3] = max h, [4] = sum pv'^2,
// [5] = sum h + sum u + sum v (finiteness, plane.cpp:176-181).  Rows in parallel, combined in row order.
template <class T>
__global__ __launch_bounds__(256) void k_plane_reduce(PlaneGeo P, const T *__restrict__ x, double *__restrict__ out) {
  const int b = blockIdx.x, t = threadIdx.x, nx = P.nx, ny = P.ny;
  const size_t nh = static_cast<size_t>(nx) * ny, nv = static_cast<size_t>(nx) * (ny + 1), per = 2 * nh + nv;
  const T *h = x + b * per, *u = h + nh, *v = u + nh;
  __shared__ double part[6][256];
  Kahan sh, se, spv, sf;
  double smax = 0, hmax = 0;
  for (int j = t; j < ny; j += blockDim.x) {
    Kahan rh, re, rpv, rf;
    for (int i = 0; i < nx; i++) {
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
    }
    if (j == ny - 1)
      for (int i = 0; i < nx; i++) rf.add(static_cast<double>(v[static_cast<size_t>(ny) * nx + i]));
    sh.add(rh.value());
    se.add(re.value());
    spv.add(rpv.value());
    sf.add(rf.value());
  }
  part[0][t] = sh.value();
  part[1][t] = se.value();
  part[2][t] = smax;
  part[3][t] = hmax;
  part[4][t] = spv.value();
  part[5][t] = sf.value();
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
    out[static_cast<size_t>(b) * 8 + t] = r;
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

uint64_t splitmix(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
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
  virtual void steps(double dt, int64_t count) = 0;
  virtual void reduce(double *out8B) = 0;  // B x 8
  virtual void tendency(int b, double *dh, double *du, double *dv) = 0;
  virtual void set_tile(int tile) = 0;
  virtual void reset() = 0;
  void sync() { LMW_HIP(hipStreamSynchronize(stream)); }
};

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
  // whole RK4 steps, one fused launch each (k_plane_step), the state ping-ponging between x0 and xa
  int tile = 0;  // fused: 0 = 32 x 32 interior, 512 threads; 1 = 32 x 16, 256 threads; 2 = 64 x 32, 512 threads; 3 = staged
  template <int TX, int TY, int NT> void launch_step(double dt) {
    constexpr int EX = TX + 8, EY = TY + 8;
    const size_t lds = (3 * (2 * static_cast<size_t>(EX) * EY + static_cast<size_t>(EX) * (EY + 1)) + static_cast<size_t>(EX) * EY +
                        static_cast<size_t>(EX) * (EY + 1)) * sizeof(T);
    dim3 grd((prm.nx + TX - 1) / TX, (prm.ny + TY - 1) / TY, B), blk(NT);
    hipLaunchKernelGGL((k_plane_step<T, TX, TY, NT>), grd, blk, lds, stream, geo(), dt, x0.p, xa.p);
    LMW_LAUNCH_CHECK();
    std::swap(x0.p, xa.p);
  }
  void set_tile(int t) override {
    if (t < 0 || t > 3) invalid("plane tile must be 0, 1, 2 or 3");
    if (t == 2 && sizeof(T) == 8) invalid("the 64 x 32 plane tile needs more LDS than a workgroup has in FP64");
    tile = t;
  }
  void steps(double dt, int64_t count) override {
    if (tile == 3) {
      steps_staged(dt, count);
      return;
    }
    for (int64_t s = 0; s < count; s++) {
      if (tile == 1) launch_step<32, 16, 256>(dt);
      else if (tile == 2) launch_step<64, 32, 512>(dt);
      else launch_step<32, 32, 512>(dt);
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
  void reduce(double *out) override {
    hipLaunchKernelGGL(k_plane_reduce<T>, dim3(B), dim3(256), 0, stream, geo(), x0.p, red.p);
    LMW_LAUNCH_CHECK();
    red.download(out, static_cast<size_t>(B) * 8, 0, stream);
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

}  // namespace lmw

using namespace lmw;

struct lmw_plane {
  std::unique_ptr<PlaneCoreBase> core;
  int precision = LMW_FP64;
};

namespace {
#define NEEDP(p) \
  if (!(p)) invalid("null pointer argument")

lmw_plane &P(lmw_plane *p) {
  NEEDP(p);
  return *p;
}
void check_member(const lmw_plane &p, int member) {
  if (member < 0 || member >= p.core->B) invalid("member index out of range");
}

// stable_dt (plane.cpp:131-141), the smallest over the members
double plane_stable_dt(lmw_plane &p, double safety) {
  PlaneCoreBase &c = *p.core;
  std::vector<double> r(static_cast<size_t>(c.B) * 8);
  if (!c.prm.linear) c.reduce(r.data());
  const double dx = c.prm.lx / c.prm.nx, dy = c.prm.ly / c.prm.ny;
  double best = INFINITY;
  for (int b = 0; b < c.B; b++) {
    const double hmax = c.prm.linear ? c.prm.depth : r[static_cast<size_t>(b) * 8 + 3];
    const double umax = c.prm.linear ? 0 : r[static_cast<size_t>(b) * 8 + 2];
    if (!(hmax > 0)) runtime("the depth must be positive to step the model");
    const double cc = std::sqrt(c.prm.gravity * hmax) + umax;
    const double inv = cc * std::sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy));
    const double visc = c.prm.viscosity * (4.0 / (dx * dx) + 4.0 / (dy * dy));
    best = std::min(best, safety * 2.0 / (inv + visc + std::fabs(c.prm.f0) + std::fabs(c.prm.beta) * c.prm.ly));
  }
  return best;
}
}  // namespace

extern "C" {

lmw_status lmw_plane_create(const lmw_plane_params *params, int batch, int precision, lmw_plane **out) {
  return guard([&] {
    NEEDP(out && params);
    *out = nullptr;
    const lmw_plane_params &p = *params;
    // plane.cpp:21-23
    if (p.nx < 4 || p.ny < 4 || p.nx > 4000 || p.ny > 4000 || static_cast<int64_t>(p.nx) * p.ny > 8000000)
      invalid("grid must be at least 4 x 4 and at most 8 million cells");
    if (!(p.lx > 0 && p.ly > 0)) invalid("domain sizes must be positive");
    if (batch < 1) invalid("a plane batch needs at least one member");
    auto h = std::make_unique<lmw_plane>();
    if (precision == LMW_FP64) h->core = std::make_unique<PlaneCore<double>>(p, batch);
    else if (precision == LMW_FP32) h->core = std::make_unique<PlaneCore<float>>(p, batch);
    else invalid("plane precision must be LMW_FP64 or LMW_FP32");
    h->precision = precision;
    *out = h.release();
  });
}
lmw_status lmw_plane_set_tile(lmw_plane *p, int tile) {
  return guard([&] { P(p).core->set_tile(tile); });
}
lmw_status lmw_plane_destroy(lmw_plane *p) {
  return guard([&] { delete p; });
}
lmw_status lmw_plane_set_params(lmw_plane *p, const lmw_plane_params *params) {
  return guard([&] {
    NEEDP(params);
    lmw_plane_params &q = P(p).core->prm;
    q.gravity = params->gravity;
    q.depth = params->depth;
    q.f0 = params->f0;
    q.beta = params->beta;
    q.viscosity = params->viscosity;
    q.drag = params->drag;
    q.linear = params->linear;
  });
}
lmw_status lmw_plane_reset(lmw_plane *p) {
  return guard([&] { P(p).core->reset(); });
}
lmw_status lmw_plane_perturb(lmw_plane *p, int member, double amplitude, uint64_t seed) {
  return guard([&] {
    lmw_plane &pl = P(p);
    check_member(pl, member);
    PlaneCoreBase &c = *pl.core;
    // plane.cpp:34-57: seeded noise, three 1-2-1 passes, zero mean, scaled to the rms -- host set-up
    const int nx = c.prm.nx, ny = c.prm.ny;
    const size_t nh = c.nh();
    uint64_t st = seed * 0x9E3779B97F4A7C15ULL + 777;
    std::vector<double> n(nh), t(nh), h(nh);
    for (auto &x : n) {
      double s = 0;
      for (int i = 0; i < 12; i++) s += static_cast<double>(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
      x = s - 6.0;
    }
    for (int pass = 0; pass < 3; pass++) {
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
          t[static_cast<size_t>(j) * nx + i] = 0.25 * n[static_cast<size_t>(j) * nx + (i + nx - 1) % nx] + 0.5 * n[static_cast<size_t>(j) * nx + i] +
                                               0.25 * n[static_cast<size_t>(j) * nx + (i + 1) % nx];
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
          n[static_cast<size_t>(j) * nx + i] = 0.25 * t[static_cast<size_t>(std::max(j - 1, 0)) * nx + i] + 0.5 * t[static_cast<size_t>(j) * nx + i] +
                                               0.25 * t[static_cast<size_t>(std::min(j + 1, ny - 1)) * nx + i];
    }
    double mean = 0, ms = 0;
    for (double x : n) mean += x;
    mean /= n.size();
    for (double &x : n) {
      x -= mean;
      ms += x * x;
    }
    const double rms = std::sqrt(ms / n.size());
    const double scale = rms > 0 ? amplitude / rms : 0;
    c.get_field(member, 0, h.data());
    for (size_t i = 0; i < nh; i++) h[i] += n[i] * scale;
    c.set_field(member, 0, h.data());
  });
}
lmw_status lmw_plane_set_field(lmw_plane *p, int member, int field, const double *g, size_t n) {
  return guard([&] {
    lmw_plane &pl = P(p);
    check_member(pl, member);
    if (field < 0 || field > 2) invalid("unknown plane field");
    const size_t want = field == 2 ? pl.core->nv() : pl.core->nh();
    if (n != want) invalid("grid size mismatch");
    NEEDP(g);
    if (field == 2) {  // set_v_from_grid: rows 0 and ny are walls (plane.cpp:314-318)
      std::vector<double> v(g, g + n);
      const int nx = pl.core->prm.nx, ny = pl.core->prm.ny;
      for (int i = 0; i < nx; i++) {
        v[i] = 0;
        v[static_cast<size_t>(ny) * nx + i] = 0;
      }
      pl.core->set_field(member, 2, v.data());
    } else {
      pl.core->set_field(member, field, g);
    }
  });
}
lmw_status lmw_plane_get_field(lmw_plane *p, int member, int field, double *out, size_t n) {
  return guard([&] {
    lmw_plane &pl = P(p);
    check_member(pl, member);
    if (field < 0 || field > 2) invalid("unknown plane field");
    const size_t want = field == 2 ? pl.core->nv() : pl.core->nh();
    if (n != want) invalid("grid size mismatch");
    NEEDP(out);
    pl.core->get_field(member, field, out);
  });
}
lmw_status lmw_plane_step(lmw_plane *p, double dt) {
  return guard([&] {
    lmw_plane &pl = P(p);
    pl.core->steps(dt, 1);
    pl.core->t += dt;
    pl.core->sync();
  });
}
lmw_status lmw_plane_steps(lmw_plane *p, double dt, int64_t count) {
  return guard([&] {
    lmw_plane &pl = P(p);
    if (count < 0) invalid("step count must not be negative");
    pl.core->steps(dt, count);
    for (int64_t i = 0; i < count; i++) pl.core->t += dt;
    pl.core->sync();
  });
}
lmw_status lmw_plane_stable_dt(lmw_plane *p, double safety, double *out) {
  return guard([&] {
    NEEDP(out);
    *out = plane_stable_dt(P(p), safety);
  });
}
// run (plane.cpp:160-189)
lmw_status lmw_plane_run(lmw_plane *p, double seconds, double dt) {
  return guard([&] {
    lmw_plane &pl = P(p);
    PlaneCoreBase &c = *pl.core;
    if (!(seconds >= 0)) invalid("run: seconds must not be negative");
    if (seconds == 0) return;
    if (!(dt > 0)) dt = plane_stable_dt(pl, 0.5);
    const double limit = plane_stable_dt(pl, std::sqrt(0.5));
    if (!(dt <= limit))
      invalid("run: dt = " + fstr(dt) + " s exceeds the RK4 stability limit of about " + fstr(limit) + " s for the present state: use a smaller dt");
    int64_t n = static_cast<int64_t>(std::ceil(seconds / dt - 1e-9));
    if (n < 1) n = 1;
    if (n > 200000000) invalid("run: too many steps; use a larger dt");
    const double hstep = seconds / static_cast<double>(n);
    std::vector<double> r(static_cast<size_t>(c.B) * 8);
    for (int64_t i = 0; i < n;) {
      const int64_t chunk = std::min<int64_t>(64 - (i & 63), n - i);
      c.steps(hstep, chunk);
      for (int64_t q = 0; q < chunk; q++) c.t += hstep;
      i += chunk;
      if (((i - 1) & 63) == 63 || i == n) {
        c.reduce(r.data());
        for (int b = 0; b < c.B; b++)
          if (!std::isfinite(r[static_cast<size_t>(b) * 8 + 5])) runtime("the integration blew up (not finite): use a smaller dt");
        const double now = plane_stable_dt(pl, std::sqrt(0.5));
        if (!(hstep <= now))
          runtime("run: dt = " + fstr(hstep) + " s exceeds the RK4 stability limit of about " + fstr(now) + " s reached at t = " + fstr(c.t) +
                  " s (the flow accelerated): use a smaller dt");
      }
    }
    c.sync();
  });
}
lmw_status lmw_plane_diag(lmw_plane *p, int member, int which, double *out) {
  return guard([&] {
    lmw_plane &pl = P(p);
    check_member(pl, member);
    NEEDP(out);
    PlaneCoreBase &c = *pl.core;
    if (which == LMW_PDIAG_TIME) {
      *out = c.t;
      return;
    }
    std::vector<double> r(static_cast<size_t>(c.B) * 8);
    c.reduce(r.data());
    const double *q = r.data() + static_cast<size_t>(member) * 8;
    const double cells = static_cast<double>(c.nh());
    switch (which) {
      case LMW_PDIAG_MEAN_DEPTH: *out = q[0] / cells; break;
      case LMW_PDIAG_ENERGY: *out = q[1] / cells; break;
      case LMW_PDIAG_MAX_SPEED: *out = q[2]; break;
      case LMW_PDIAG_PV_ANOMALY_RMS: *out = std::sqrt(q[4] / (static_cast<double>(c.prm.ny - 1) * c.prm.nx)); break;
      default: invalid("unknown plane diagnostic");
    }
  });
}
lmw_status lmw_plane_tendency(lmw_plane *p, int member, double *dh, double *du, double *dv) {
  return guard([&] {
    lmw_plane &pl = P(p);
    check_member(pl, member);
    NEEDP(dh && du && dv);
    pl.core->tendency(member, dh, du, dv);
  });
}

}  // extern "C"
