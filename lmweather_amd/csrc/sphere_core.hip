// Spectral-transform shallow water on the sphere for gfx950: the reference's Sphere::tendencies / step
// (cpp/sphere.cpp:338-436) as three kernels per RK4 stage, batched over the members of one truncation:
//
//   k_synth      inverse Legendre: per (m, 64 northern latitudes, 4 members) the even/odd-n partial sums of
//                zeta P, phi P, psi P, chi P, psi H, chi H (sphere.cpp:150-162,179-206) folded by equatorial
//                symmetry into the Fourier coefficients of U, V, zeta, phi at both latitudes of the pair.
//   k_tend_grid  per (latitude pair, member) in LDS: two inverse complex FFTs per latitude (two real rows packed as
//                one complex row), the grid products eta = zeta + f, A = U eta, B = V eta, U phi, V phi,
//                phi + g hs + E (sphere.cpp:348-359), five forward FFTs (sphere.cpp:366-370) and the operands of
//                the analysis with the Gaussian weights and the i m factors applied (sphere.cpp:371-376).
//   k_anal       forward Legendre: per (m, 64 output rows, 4 members) the sums over northern latitudes of
//                P X + H Y (sphere.cpp:377-384), then the RK4 stage update in the epilogue (sphere.cpp:416-434).
//
// FP64 throughout.  The FFT twiddles are exp(-2 pi i q / K) rounded once on the host.
#include <cmath>

#include "fft_lds.h"
#include "sphere_core.h"
#include "sphere_epilogue.h"

namespace lmw {

const int kFp64MB[kFp64Tiles] = {2, 4, 8};  // members per block of the FP64 Legendre kernels (tile variants)

namespace {
constexpr int CH = 32;     // latitudes staged per LDS chunk in k_anal
constexpr int MAXM = 172;  // truncation <= 170 => at most 171 m values

__device__ inline int brev(int k, int logK) { return static_cast<int>(__brev(static_cast<unsigned>(k)) >> (32 - logK)); }

// eps(n, m) of sphere.cpp:11
__host__ __device__ inline double eps_nm(int n, int m) {
  return sqrt(static_cast<double>(n * n - m * m) / static_cast<double>(4 * n * n - 1));
}

// Iterative radix-2 DIT FFT over nf arrays of length L in LDS, data already in bit-reversed order.
// tw[q] = exp(-2 pi i q / (L << tws)); sign +1 uses the conjugate twiddles (backward transform, unnormalized).
__device__ void fft_lds(double2 *buf, int nf, int L, int logL, const double2 *tw, int tws, int sign) {
  const int halfL = L >> 1;
  for (int s = 1; s <= logL; s++) {
    const int half = 1 << (s - 1), tstride = (L >> s) << tws;
    for (int t = threadIdx.x; t < nf * halfL; t += blockDim.x) {
      const int f = t >> (logL - 1), bt = t & (halfL - 1);
      const int kk = bt & (half - 1), grp = bt >> (s - 1);
      const int i0 = f * L + grp * 2 * half + kk, i1 = i0 + half;
      double2 wv = tw[kk * tstride];
      if (sign > 0) wv.y = -wv.y;
      const double2 u = buf[i0], v = cmul(buf[i1], wv);
      buf[i0] = cadd(u, v);
      buf[i1] = csub(u, v);
    }
    __syncthreads();
  }
}

// Real rows of length K = 2L are transformed one row per complex array of length L: the even samples are the real
// parts and the odd samples the imaginary parts, so every row keeps the rounding of its own magnitude (as the
// reference's separate transform of each row does, sphere.cpp:123-136).  tw[q] = exp(-2 pi i q / K), q < L.
//
// Forward: arr = FFT_L(x[2k] + i x[2k+1]); returns X_m / K for 0 <= m < L (fourier_forward: a[m] / nlon).
template <class F, class C> __device__ inline C rfft_coef(const C *arr, int m, const C *tw, typename real_of<C>::type invK) {
  using T = typename real_of<C>::type;
  const T half = static_cast<T>(0.5);
  const C Zm = arr[F::phys(F::pos(m))], Zc = arr[F::phys(F::pos((F::L - m) & (F::L - 1)))];
  const C Fe = C{half * (Zm.x + Zc.x), half * (Zm.y - Zc.y)};  // (Zm + conj Zc) / 2: even samples
  const C Fo = C{half * (Zm.y + Zc.y), half * (Zc.x - Zm.x)};  // (Zm - conj Zc) / 2i: odd samples
  return cscale(cadd(Fe, cmul(tw[m], Fo)), invK);
}
// 2 K rfft_coef, formed without the halvings: (Zm + conj Zc) + tw[m] (Zm - conj Zc) / i.  Each scaling by a power of two
// is exact, so X scale(2 K rfft_coef) rounds once like rfft_coef times the same weight: scale = weight / (2K).
template <class F, class C> __device__ inline C rfft_coef2K(const C *arr, int m, const C *tw) {
  const C Zm = arr[F::phys(F::pos(m))], Zc = arr[F::phys(F::pos((F::L - m) & (F::L - 1)))];
  const C Se = C{Zm.x + Zc.x, Zm.y - Zc.y}, So = C{Zm.y + Zc.y, Zc.x - Zm.x};
  return cadd(Se, cmul(tw[m], So));
}

// Backward: entry m' (0 <= m' < L) of the half-length spectrum whose backward transform is x[2k] + i x[2k+1] for
// the real row x_k = sum_m a_m exp(2 pi i m k / K) with a_0 = Re F_0, a_m = F_m, a_{K-m} = conj F_m (m = 1..N) and
// zero elsewhere (fourier_backward, sphere.cpp:130-136).  F_m is read at src[4 m].
template <class TS, class C> __device__ inline C irfft_entry(const TS *__restrict__ src, int mp, int L, int N, const C *tw) {
  C lo = mkc<C>(0, 0), hi = lo;  // a_{m'} and a_{m'+L}
  if (mp == 0) lo = C{to_c<C>(src[0]).x, 0};
  else if (mp <= N) lo = to_c<C>(src[4 * mp]);
  if (mp >= 1 && L - mp <= N) hi = cconj(to_c<C>(src[4 * (L - mp)]));
  const C E = cadd(lo, hi), O = cmul(csub(lo, hi), cconj(tw[mp]));
  return C{E.x - O.y, E.y + O.x};  // E + i O
}

// Writes the analysis operands of one (member, m, latitude pair) from the per-latitude values (north n, south s).
// Layout per (b, m): [parity p][part: 0 = P operand X, 1 = H operand Y][jh][4], where even rows (p = 0) use the
// north/south sums X+ = Xn + Xs and differences Y- = Yn - Ys, odd rows X- and Y+ (P(-mu) = (-1)^(n-m) P(mu),
// H(-mu) = -(-1)^(n-m) H(mu)).
template <class TS, class C>
__device__ inline void store_operands(TS *Gb, size_t bm, int Jh, int jh, const C Xn[4], const C Xs[4], const C Yn[4], const C Ys[4]) {
  TS *o = Gb + bm * 16 * Jh + static_cast<size_t>(jh) * 4;
  const size_t part = static_cast<size_t>(Jh) * 4;
  if constexpr (sizeof(TS) == 8) {  // FP32 records: 32 bytes, 32-byte aligned, written as two 16-byte stores each
    TS v[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      v[0][i] = to_c<TS>(cadd(Xn[i], Xs[i]));
      v[1][i] = to_c<TS>(csub(Yn[i], Ys[i]));
      v[2][i] = to_c<TS>(csub(Xn[i], Xs[i]));
      v[3][i] = to_c<TS>(cadd(Yn[i], Ys[i]));
    }
#pragma unroll
    for (int p = 0; p < 4; p++) {
      float4 *o4 = reinterpret_cast<float4 *>(o + p * part);
      o4[0] = make_float4(v[p][0].x, v[p][0].y, v[p][1].x, v[p][1].y);
      o4[1] = make_float4(v[p][2].x, v[p][2].y, v[p][3].x, v[p][3].y);
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    o[i] = to_c<TS>(cadd(Xn[i], Xs[i]));             // p 0, X
    o[part + i] = to_c<TS>(csub(Yn[i], Ys[i]));      // p 0, Y
    o[2 * part + i] = to_c<TS>(csub(Xn[i], Xs[i]));  // p 1, X
    o[3 * part + i] = to_c<TS>(cadd(Yn[i], Ys[i]));  // p 1, Y
  }
}
}  // namespace

// ------------------------------------------------------------------------------------------- set-up kernels

// Gauss-Legendre nodes and weights (sphere.cpp:47-64), one thread per root pair.
__global__ void k_gauss(int n, double *x, double *w) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (n + 1) / 2) return;
  const double pi = 3.14159265358979323846;
  double z = cos(pi * (i + 0.75) / (n + 0.5)), pp = 1;
  for (int it = 0; it < 100; it++) {
    double p1 = 1, p2 = 0;
    for (int j = 1; j <= n; j++) {
      double p3 = p2;
      p2 = p1;
      p1 = ((2.0 * j - 1) * z * p2 - (j - 1.0) * p3) / j;
    }
    pp = n * (z * p1 - p2) / (z * z - 1);
    double dz = p1 / pp;
    z -= dz;
    if (fabs(dz) < 1e-15) break;
  }
  x[i] = -z;
  x[n - 1 - i] = z;
  w[i] = w[n - 1 - i] = 2 / ((1 - z * z) * pp * pp);
}

// Orthonormal P_n^m and H_n^m = (1 - mu^2) dP/dmu at the northern latitudes (sphere.cpp:66-73 and :84-96),
// one thread per (m, latitude), the three-term recurrence kept in registers.
__global__ void k_tables(Geo G, const double *mu, double *Ps, double *Hs, double *Pa, double *Ha) {
#pragma clang fp contract(off)
  const int m = blockIdx.x, jh = threadIdx.x;
  if (jh >= G.Jh) return;
  const double x = mu[G.Jh + jh];
  const int kpe = G.kpe[m], kp = kpe + G.kpo[m];
  const size_t base = static_cast<size_t>(G.poff[m]);
  const double c = sqrt(fmax(0.0, 1 - x * x));
  double p = 1 / sqrt(2.0);
  for (int k = 1; k <= m; k++) p *= sqrt((2.0 * k + 1) / (2.0 * k)) * c;
  double prev2 = 0.0, prev = p, cur = sqrt(2.0 * m + 3) * x * p;  // P_{n-1}, P_n, P_{n+1} at n = m
  for (int n = m; n <= G.N; n++) {
    const double minus = (n > m) ? prev2 : 0.0;
    const double H = -n * eps_nm(n + 1, m) * cur + (n + 1) * eps_nm(n, m) * minus;
    const int k = n - m, r = (k & 1) ? kpe + (k >> 1) : (k >> 1);
    Ps[(base + r) * G.Jhp + jh] = prev;
    Hs[(base + r) * G.Jhp + jh] = H;
    Pa[base * G.Jhp + static_cast<size_t>(jh) * kp + r] = prev;
    Ha[base * G.Jhp + static_cast<size_t>(jh) * kp + r] = H;
    const double next = (n + 2 <= G.N + 1) ? (x * cur - eps_nm(n + 1, m) * prev) / eps_nm(n + 2, m) : 0.0;
    prev2 = prev;
    prev = cur;
    cur = next;
  }
}

// f = 2 Omega (-cos(lon) cos(lat) sin(alpha) + mu cos(alpha)) (sphere.cpp:111-118) for one member.
__global__ void k_coriolis(Geo G, const double *mu, double *f, double alpha) {
#pragma clang fp contract(off)
  const int j = blockIdx.x;
  const double pi = 3.14159265358979323846;
  const double c = sqrt(1 - mu[j] * mu[j]);
  for (int k = threadIdx.x; k < G.K; k += blockDim.x) {
    const double lam = 2 * pi * k / G.K;
    f[static_cast<size_t>(j) * G.K + k] = 2 * G.omega * (-cos(lam) * c * sin(alpha) + mu[j] * cos(alpha));
  }
}

// Analytic initial states of Williamson et al. (1992, J. Comput. Phys. 102:211-224) at the Gauss grid: the set-up
// operators of cpp/main.cpp:174-291.  in3[local member][h, u, v][j][k] for the members of blockIdx.y; test case 5 also
// writes its conical mountain into the orography plane hs_plane (when given), and SETUP_TRACER writes test case 1's
// cosine bell into plane 0 of in3.
// lat = asin(mu_j), lon = 2 pi k / K (sphere.cpp:120-121).
__global__ void k_setup(Geo G, int which, double alpha, const double *__restrict__ mu, double *__restrict__ in3,
                        double *__restrict__ hs_plane) {
  const int j = blockIdx.x, bl = blockIdx.y, K = G.K;
  const size_t plane = static_cast<size_t>(G.J) * K, row = static_cast<size_t>(j) * K;
  const double pi = 3.14159265358979323846, a = G.a, om = G.omega, g = G.g;
  const double lat = asin(mu[j]), cl = cos(lat), sl = sin(lat);
  double *h = in3 + static_cast<size_t>(bl) * 3 * plane + row, *u = h + plane, *v = u + plane;
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    const double lon = 2 * pi * k / K;
    if (which == LMW_SETUP_WILLIAMSON2 || which == LMW_SETUP_WILLIAMSON1 || which == SETUP_TRACER) {
      // solid-body rotation about an axis tilted by alpha, once round in 12 days (test cases 1 and 2)
      const double u0 = 2 * pi * a / (12 * 86400.0);
      if (which == SETUP_TRACER) {  // cosine bell of 1000 m, radius a/3, centred at (270 E, 0 N)
        const double R = a / 3, lonc = 1.5 * pi;
        const double c = cos(lat) * cos(lon - lonc);  // sin(latc) = 0
        const double rr = a * acos(fmax(-1.0, fmin(1.0, c)));
        h[k] = rr < R ? 500.0 * (1 + cos(pi * rr / R)) : 0.0;
        continue;
      }
      u[k] = u0 * (cl * cos(alpha) + sl * cos(lon) * sin(alpha));
      v[k] = -u0 * sin(lon) * sin(alpha);
      if (which == LMW_SETUP_WILLIAMSON1) {
        h[k] = 1000.0;
      } else {  // geostrophically balanced height, g h0 = 2.94e4 m2/s2
        const double t = -cos(lon) * cl * sin(alpha) + sl * cos(alpha);
        h[k] = 2.94e4 / g - (a * om * u0 + 0.5 * u0 * u0) * t * t / g;
      }
    } else if (which == LMW_SETUP_WILLIAMSON5) {
      // test case 2's zonal flow (u0 = 20 m/s, h0 = 5960 m) over a 2000 m cone of radius pi/9 at (270 E, 30 N); the
      // fluid depth is the balanced free surface minus the mountain
      const double u0 = 20.0, R = pi / 9, lonc = 1.5 * pi, latc = pi / 6;
      const double rr = sqrt(fmin(R * R, (lon - lonc) * (lon - lonc) + (lat - latc) * (lat - latc)));
      const double mount = 2000.0 * (1 - rr / R);
      if (hs_plane && bl == 0) hs_plane[row + k] = mount;  // one plane, shared by the members of the launch
      u[k] = u0 * cl;
      v[k] = 0.0;
      h[k] = 5960.0 - (a * om * u0 + 0.5 * u0 * u0) * sl * sl / g - mount;
    } else {  // test case 6: Rossby-Haurwitz wave of wavenumber R = 4, w = K = 7.848e-6 1/s, h0 = 8000 m
      const double w = 7.848e-6, Kc = 7.848e-6, R = 4;
      const double c2 = cl * cl, cR1 = pow(cl, R - 1), cR = pow(cl, R), c2R = pow(cl, 2 * R);
      u[k] = a * w * cl + a * Kc * cR1 * (R * sl * sl - c2) * cos(R * lon);
      v[k] = -a * Kc * R * cR1 * sl * sin(R * lon);
      const double A = w / 2 * (2 * om + w) * c2 + Kc * Kc / 4 * c2R * ((R + 1) * c2 + (2 * R * R - R - 2) - 2 * R * R / c2);
      const double B = 2 * (om + w) * Kc / ((R + 1) * (R + 2)) * cR * ((R * R + 2 * R + 2) - (R + 1) * (R + 1) * c2);
      const double C = Kc * Kc / 4 * c2R * ((R + 1) * c2 - (R + 2));
      h[k] = 8000.0 + a * a / g * (A + B * cos(R * lon) + C * cos(2 * R * lon));
    }
  }
}

// ------------------------------------------------------------------------------------------- the step kernels

// Inverse Legendre transform of (zeta, delta, phi) to the Fourier coefficients of U, V, zeta, phi.  MB members per
// block (one per threadIdx.y).
template <int MB>
__global__ __launch_bounds__(64 * MB) void k_synth(Geo G, const double *__restrict__ Ps, const double *__restrict__ Hs,
                                               const double2 *__restrict__ spec, double2 *__restrict__ Fs, int b0, int nb) {
  const int m = blockIdx.x, tx = threadIdx.x, ty = threadIdx.y;
  const int jh = blockIdx.y * 64 + tx;
  const int bl = blockIdx.z * MB + ty, b = b0 + bl;
  const int Km = G.N - m + 1, Ke = (Km + 1) >> 1, Ko = Km >> 1;
  __shared__ double2 cz[MB][MAXM], cf[MB][MAXM], cp[MB][MAXM], cc[MB][MAXM];
  const double a = G.a;
  if (bl < nb) {
    const size_t off = static_cast<size_t>(G.moff[m]);
    for (int k = tx; k < Km; k += 64) {
      const int n = m + k;
      const double kn = n >= 1 ? -a * a / (n * (n + 1.0)) : 0.0;  // psi, chi = -a^2/(n(n+1)) (zeta, delta): :185-190
      const double2 z = spec[G.at(b, 0, off + k)], d = spec[G.at(b, 1, off + k)];
      cz[ty][k] = z;
      cf[ty][k] = spec[G.at(b, 2, off + k)];
      cp[ty][k] = cscale(z, kn);
      cc[ty][k] = cscale(d, kn);
    }
  }
  __syncthreads();
  if (bl >= nb || jh >= G.Jh) return;
  double2 zE = cx(0, 0), fE = zE, pE = zE, cE = zE, pHE = zE, cHE = zE;
  double2 zO = zE, fO = zE, pO = zE, cO = zE, pHO = zE, cHO = zE;
  const size_t col = static_cast<size_t>(G.poff[m]) * G.Jhp + jh;
  const double *pc = Ps + col, *hc = Hs + col;
  const int kpe = G.kpe[m];
  for (int r = 0; r < Ke; r++) {
    const double P = pc[static_cast<size_t>(r) * G.Jhp], H = hc[static_cast<size_t>(r) * G.Jhp];
    const int k = 2 * r;
    cfma(zE, cz[ty][k], P);
    cfma(fE, cf[ty][k], P);
    cfma(pE, cp[ty][k], P);
    cfma(cE, cc[ty][k], P);
    cfma(pHE, cp[ty][k], H);
    cfma(cHE, cc[ty][k], H);
  }
  for (int r = 0; r < Ko; r++) {
    const double P = pc[static_cast<size_t>(kpe + r) * G.Jhp], H = hc[static_cast<size_t>(kpe + r) * G.Jhp];
    const int k = 2 * r + 1;
    cfma(zO, cz[ty][k], P);
    cfma(fO, cf[ty][k], P);
    cfma(pO, cp[ty][k], P);
    cfma(cO, cc[ty][k], P);
    cfma(pHO, cp[ty][k], H);
    cfma(cHO, cc[ty][k], H);
  }
  const double dm = m;
  for (int s = 0; s < 2; s++) {
    const double sg = s == 0 ? 1.0 : -1.0;  // P even in mu for even n-m; H the opposite
    const int j = s == 0 ? G.Jh + jh : G.Jh - 1 - jh;
    const double2 Z = cadd(zE, cscale(zO, sg)), F = cadd(fE, cscale(fO, sg));
    const double2 psP = cadd(pE, cscale(pO, sg)), chP = cadd(cE, cscale(cO, sg));
    const double2 psH = cadd(cscale(pHE, sg), pHO), chH = cadd(cscale(cHE, sg), cHO);
    // U_m = (i m chi P - psi H) / a, V_m = (i m psi P + chi H) / a  (sphere.cpp:200-201)
    const double2 u = csub(cim(dm, chP), psH), v = cadd(cim(dm, psP), chH);
    double2 *o = Fs + ((static_cast<size_t>(b) * G.J + j) * G.M1 + m) * 4;
    o[0] = cx(u.x / a, u.y / a);
    o[1] = cx(v.x / a, v.y / a);
    o[2] = Z;
    o[3] = F;
  }
}

// Grid-point work of one latitude pair: inverse FFTs, nonlinear products, forward FFTs, analysis operands.
// Rows in LDS: inverse U, V, zeta, phi and forward A, B, U phi, V phi, phi + g hs + E, each at (north, south).
// Storage TS of the Fourier coefficients and analysis operands: double2 (FP64 path) or float2 (mixed precision).
// mflags[b]: bit 0 = rotation axis untilted (f = 2 Omega mu computed, not read), bit 1 = no orography (hs = 0),
// bits 2.. = the member's plane of hs (members with the same orography share one).
// Grid-point part of one RK4 stage for one latitude pair and member: inverse FFTs of U, V, zeta, phi (north and
// south rows), the nonlinear products (sphere.cpp:348-358), forward FFTs of the five products and the analysis
// operands.  All global loads a thread needs (coefficients, twiddles, orography, f) are issued before the first
// result is used.  TS is the storage type of the Fourier rows and operands; the arithmetic runs in its precision
// (double2: the FP64 step; float2: the mixed-precision step, whose rows hold FP32 values already).
template <int R1, int R2, class TS>
__global__ __launch_bounds__(256) void k_tend_grid(Geo G, const TS *__restrict__ Fs, const double *__restrict__ fgrid,
                                                   const double *__restrict__ hs, const TS *__restrict__ twg,
                                                   const double *__restrict__ mu, const double *__restrict__ w,
                                                   TS *__restrict__ Gb, int b0, const int *__restrict__ mflags) {
  using F = Fft2<R1, R2>;
  using C = TS;
  using T = typename real_of<C>::type;
  constexpr int L = F::L, P = F::P;
  constexpr int NIT = (2 * L + 255) / 256;  // blockDim.x == 256
  extern __shared__ double2 lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  const int K = 2 * L, M1 = G.M1, tid = threadIdx.x;
  // XCD-aware order: the hardware deals consecutive blocks to the 8 XCDs in turn, so block x of a member takes latitude
  // pair (x mod 8) Jh/8 + x/8 and each XCD works through a contiguous band of latitude pairs.  The 32-byte operand
  // records of neighbouring pairs share cache lines, which then fill in one XCD's L2 before they are written back.
  const int jh = (G.Jh & 7) ? static_cast<int>(blockIdx.x) : static_cast<int>((blockIdx.x & 7) * (G.Jh >> 3) + (blockIdx.x >> 3));
  const int b = b0 + blockIdx.y;
  const int jn = G.Jh + jh, js = G.Jh - 1 - jh;
  C *buf = lds, *tw = lds + 10 * P;
  const int k = tid;  // one pair of longitudes (2k, 2k+1) per thread in the grid products: L <= 256
  const int flags = mflags[b];
  const bool untilted = flags & 1, flat = flags & 2;
  const size_t hplane = static_cast<size_t>(flags >> 2);  // this member's orography plane
  // Every global load of the prologue is issued before the first value is used, from a clamped valid address where
  // a lane has nothing to load (its value is zeroed afterwards), so the loads are in flight together and the thread
  // waits once: orography and f of this thread's longitudes of both rows, the four fields' coefficients at m' and
  // L - m' of each inverse-FFT input, the twiddles.  Orography and f are loaded whatever the member's flags and
  // selected only where they are used (loaded inside a branch on the flags, they were converted inside it, which
  // waited for them at once); a member without orography or with an untilted axis reads the first row of plane 0
  // instead (every group holds at least one plane), which all its blocks share from the cache.
  const int kc = k < L ? k : L - 1;
  double hraw[2][2], fraw[2][2];
#pragma unroll
  for (int s = 0; s < 2; s++) {
    const size_t hrow = flat ? 0 : hplane * G.J + (s ? js : jn);
    const size_t frow = untilted ? 0 : static_cast<size_t>(b) * G.J + (s ? js : jn);
#pragma unroll
    for (int e = 0; e < 2; e++) {
      hraw[s][e] = hs[hrow * K + 2 * kc + e];
      fraw[s][e] = fgrid[frow * K + 2 * kc + e];
    }
  }
  // inverse real FFT inputs (the arithmetic of irfft_entry): per (row s, m') the four fields' coefficients at m'
  // and L - m', 4 contiguous complex values each
  {
    TS lo[NIT][4], hi[NIT][4];
    C twm[NIT];
#pragma unroll
    for (int q = 0; q < NIT; q++) {
      const int it = min(tid + 256 * q, 2 * L - 1), s = it / L, mp = it - s * L;
      const TS *row = Fs + (static_cast<size_t>(b) * G.J + (s ? js : jn)) * M1 * 4;
      const int ml = min(mp, G.N), mh = min(L - mp, G.N);  // (clamped: lanes without a coefficient read a valid one)
#pragma unroll
      for (int f = 0; f < 4; f++) lo[q][f] = row[4 * ml + f];
#pragma unroll
      for (int f = 0; f < 4; f++) hi[q][f] = row[4 * mh + f];
      twm[q] = to_c<C>(twg[mp]);
    }
    for (int i = tid; i < K; i += blockDim.x) tw[i] = to_c<C>(twg[i]);
#pragma unroll
    for (int q = 0; q < NIT; q++) {
      const int it = tid + 256 * q, s = it / L, mp = it - s * L;
      if (it < 2 * L) {
        const bool has_lo = mp <= G.N, has_hi = mp >= 1 && L - mp <= G.N;
#pragma unroll
        for (int f = 0; f < 4; f++) {
          C a = has_lo ? to_c<C>(lo[q][f]) : mkc<C>(0, 0);
          if (mp == 0) a.y = 0;
          const C c = has_hi ? cconj(to_c<C>(hi[q][f])) : mkc<C>(0, 0);
          const C E = cadd(a, c), O = cmul(csub(a, c), cconj(twm[q]));
          buf[(2 * f + s) * P + F::phys(mp)] = C{E.x - O.y, E.y + O.x};  // E + i O
        }
      }
    }
  }
  T hsv[2][2], fv[2][2];
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int e = 0; e < 2; e++) {
      hsv[s][e] = k < L ? static_cast<T>(hraw[s][e]) : T(0);
      fv[s][e] = k < L ? static_cast<T>(fraw[s][e]) : T(0);
    }
  __syncthreads();
  fft2<R1, R2, +1>(buf, 8, tw);
  // mixed precision: the phi rows hold phi' = phi - phibar (phibar = phi_00 / sqrt 2, the area mean), so the phi
  // fluxes and the Bernoulli function are formed with phi'; the forward transform's epilogue adds the mean's exact
  // contribution -phibar delta (the mean's part of the Bernoulli function only reaches n = 0, whose tendency is 0)
  const T om = static_cast<T>(1 - mu[jn] * mu[jn]);  // identical for the southern row: mu(s) = -mu(n)
  // FP32 rows: E = (u^2 + v^2) times 1 / (2 om), rounded once from FP64, instead of four FP32 divisions per thread
  const T r2om = static_cast<T>(0.5 / (1 - mu[jn] * mu[jn]));
  const T g = static_cast<T>(G.g);
  C res[10];
  if (k < L) {
    const int pk = F::phys(F::pos(k));
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const int j = s ? js : jn;
      const C U = buf[(0 + s) * P + pk], V = buf[(2 + s) * P + pk], Z = buf[(4 + s) * P + pk], Ph = buf[(6 + s) * P + pk];
      const T fj = static_cast<T>((2 * G.omega) * mu[j]);
      T o[5][2];
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const T u = e ? U.y : U.x, v = e ? V.y : V.x, z = e ? Z.y : Z.x, ph = e ? Ph.y : Ph.x;
        // untilted f = 2 Omega (-cos(lon) cos(lat) 0 + mu 1) = (2 Omega) mu exactly (sphere.cpp:116)
        const T eta = z + (untilted ? fj : fv[s][e]);
        o[0][e] = u * eta;
        o[1][e] = v * eta;
        o[2][e] = u * ph;
        o[3][e] = v * ph;
        const T ke = u * u + v * v;
        o[4][e] = ph + g * (flat ? T(0) : hsv[s][e]) + (sizeof(T) == 4 ? ke * r2om : ke / (2 * om));  // g (h + hs) + E (sphere.cpp:357)
      }
#pragma unroll
      for (int q = 0; q < 5; q++) res[2 * q + s] = C{o[q][0], o[q][1]};
    }
  }
  __syncthreads();
  if (k < L) {
    const int pk = F::phys(k);
#pragma unroll
    for (int r = 0; r < 10; r++) buf[r * P + pk] = res[r];
  }
  __syncthreads();
  fft2<R1, R2, -1>(buf, 10, tw);
  const T wj = static_cast<T>(w[jn]), wd = static_cast<T>(w[jn] / (1 - mu[jn] * mu[jn]) / G.a), invK = static_cast<T>(1.0 / K);
  for (int m = tid; m < M1; m += blockDim.x) {
    if constexpr (sizeof(TS) == 8) {
      // FP32: one 80-byte record per (member, m, latitude pair) of the five Fourier coefficients at both latitudes,
      // weighted (wd A, wd B, wd U phi, wd V phi, wj E; north then south); the forward transform forms the i m
      // factors and the north/south combinations of each row parity as it loads them (five stored values per
      // latitude instead of eight operands per parity).  The weights absorb 1 / (2K): the same bits as the weight
      // times rfft_coef.
      const T wd2 = wd * (invK * T(0.5)), wj2 = wj * (invK * T(0.5));
      C c[10];
#pragma unroll
      for (int r = 0; r < 10; r++) c[r] = rfft_coef2K<F>(buf + r * P, m, tw);
      float4 *o = reinterpret_cast<float4 *>(Gb + ((static_cast<size_t>(b) * M1 + m) * G.Jh + jh) * 10);
      const C v[10] = {cscale(c[0], wd2), cscale(c[2], wd2), cscale(c[4], wd2), cscale(c[6], wd2), cscale(c[8], wj2),
                       cscale(c[1], wd2), cscale(c[3], wd2), cscale(c[5], wd2), cscale(c[7], wd2), cscale(c[9], wj2)};
#pragma unroll
      for (int i = 0; i < 5; i++) o[i] = make_float4(v[2 * i].x, v[2 * i].y, v[2 * i + 1].x, v[2 * i + 1].y);
      continue;
    }
    C c[10];
#pragma unroll
    for (int r = 0; r < 10; r++) c[r] = rfft_coef<F>(buf + r * P, m, tw, invK);
    const T dm = static_cast<T>(m);
    // dzeta -= (i m A wd) P - (B wd) H; ddelta += (i m B wd) P + (A wd) H + (E w) P n(n+1)/a^2;
    // dphi -= (i m U phi wd) P - (V phi wd) H   (sphere.cpp:374-383); index 0 north, 1 south
    C X[2][4], Y[2][4];
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const C A = c[0 + s], B = c[2 + s], UP = c[4 + s], VP = c[6 + s], PE = c[8 + s];
      X[s][0] = cscale(cim(dm, A), -wd);
      Y[s][0] = cscale(B, wd);
      X[s][1] = cscale(cim(dm, B), wd);
      Y[s][1] = cscale(A, wd);
      X[s][2] = cscale(PE, wj);
      Y[s][2] = mkc<C>(0, 0);
      X[s][3] = cscale(cim(dm, UP), -wd);
      Y[s][3] = cscale(VP, wd);
    }
    store_operands(Gb, static_cast<size_t>(b) * M1 + m, G.Jh, jh, X[0], X[1], Y[0], Y[1]);
  }
}

// Grid inputs to analysis operands: the set_state_grid analyses (sphere.cpp:217-231), analyze_divergence
// (:164-177) and analyze_scalar (:138-149).  Rows: STATE U, V, phi; DIV A, B; SCALAR the field; each (north, south).
template <int R1, int R2>
__global__ __launch_bounds__(256) void k_prep(Geo G, int mode, const double *__restrict__ in3, const double2 *__restrict__ twg,
                                              const double *__restrict__ mu, const double *__restrict__ w, double2 *__restrict__ Gb, int b0) {
  using F = Fft2<R1, R2>;
  constexpr int L = F::L, P = F::P;
  extern __shared__ double2 lds[];
  const int K = 2 * L, M1 = G.M1, tid = threadIdx.x;
  const int jh = blockIdx.x, b = b0 + blockIdx.y;
  const int jn = G.Jh + jh, js = G.Jh - 1 - jh;
  double2 *buf = lds, *tw = lds + 10 * P;
  const size_t plane = static_cast<size_t>(G.J) * K;
  const double *g0 = in3 + static_cast<size_t>(blockIdx.y) * 3 * plane;
  for (int i = tid; i < K; i += blockDim.x) tw[i] = twg[i];
  const double c = sqrt(1 - mu[jn] * mu[jn]);
  const int nrows = (mode == PREP_STATE || mode == PREP_TRACER) ? 6 : mode == PREP_DIV ? 4 : 2;
  for (int idx = tid; idx < nrows * L; idx += blockDim.x) {
    const int r = idx / L, k = idx - r * L;
    const int field = r >> 1;
    const size_t o = static_cast<size_t>((r & 1) ? js : jn) * K + 2 * k;
    double x0, x1;
    if (mode == PREP_STATE) {  // U = u cos, V = v cos, phi = g h (sphere.cpp:221-229); in3 = h, u, v
      const double *src = g0 + (field == 2 ? 0 : (field + 1)) * plane;
      const double f = field == 2 ? G.g : c;
      x0 = field == 2 ? f * src[o] : src[o] * f;
      x1 = field == 2 ? f * src[o + 1] : src[o + 1] * f;
    } else {
      const double *src = g0 + field * plane;
      x0 = src[o];
      x1 = src[o + 1];
    }
    buf[r * P + F::phys(k)] = cx(x0, x1);
  }
  __syncthreads();
  fft2<R1, R2, -1>(buf, nrows, tw);
  const double wj = w[jn], wd = w[jn] / (1 - mu[jn] * mu[jn]) / G.a, invK = 1.0 / K;
  for (int m = tid; m < M1; m += blockDim.x) {
    const double dm = m;
    double2 X[2][4], Y[2][4];
    for (int s = 0; s < 2; s++)
      for (int i = 0; i < 4; i++) X[s][i] = Y[s][i] = cx(0, 0);
    for (int s = 0; s < 2; s++) {
      if (mode == PREP_STATE) {
        const double2 U = rfft_coef<F>(buf + (0 + s) * P, m, tw, invK), V = rfft_coef<F>(buf + (2 + s) * P, m, tw, invK);
        const double2 Ph = rfft_coef<F>(buf + (4 + s) * P, m, tw, invK);
        // zeta = D(V, -U), delta = D(U, V) (sphere.cpp:211-212), phi by analyze_scalar
        X[s][0] = cscale(cim(dm, V), wd);
        Y[s][0] = cscale(U, wd);
        X[s][1] = cscale(cim(dm, U), wd);
        Y[s][1] = cscale(V, -wd);
        X[s][3] = cscale(Ph, wj);
      } else if (mode == PREP_DIV) {
        const double2 A = rfft_coef<F>(buf + (0 + s) * P, m, tw, invK), B = rfft_coef<F>(buf + (2 + s) * P, m, tw, invK);
        X[s][0] = cscale(cim(dm, A), wd);
        Y[s][0] = cscale(B, -wd);
      } else if (mode == PREP_TRACER) {  // flux D(A, B) into o1, analyze_scalar(C) into o4
        const double2 A = rfft_coef<F>(buf + (0 + s) * P, m, tw, invK), B = rfft_coef<F>(buf + (2 + s) * P, m, tw, invK);
        const double2 Cq = rfft_coef<F>(buf + (4 + s) * P, m, tw, invK);
        X[s][0] = cscale(cim(dm, A), wd);
        Y[s][0] = cscale(B, -wd);
        X[s][3] = cscale(Cq, wj);
      } else {
        X[s][3] = cscale(rfft_coef<F>(buf + s * P, m, tw, invK), wj);
      }
    }
    store_operands(Gb, static_cast<size_t>(b) * M1 + m, G.Jh, jh, X[0], X[1], Y[0], Y[1]);
  }
}

// Forward Legendre transform and its epilogue (RK4 stage update, state write or raw output).  MB members per block.
template <int MB>
__global__ __launch_bounds__(64 * MB) void k_anal(Geo G, const double *__restrict__ Pa, const double *__restrict__ Ha,
                                              const double2 *__restrict__ Gb, int mode, int stage, double dt, double nu4,
                                              double2 *__restrict__ x0, double2 *__restrict__ xs, double2 *__restrict__ acc,
                                              double2 *__restrict__ raw, int b0, int nb) {
  const int m = blockIdx.x, tx = threadIdx.x, ty = threadIdx.y;
  const int r = blockIdx.y * 64 + tx;
  const int bl = blockIdx.z * MB + ty, b = b0 + bl;
  const int kpe = G.kpe[m], kp = kpe + G.kpo[m], Km = G.N - m + 1;
  const size_t q = static_cast<size_t>(G.moff[m]);  // + k below
  const bool odd = r >= kpe;
  const int k = odd ? 2 * (r - kpe) + 1 : 2 * r;
  const bool valid = (bl < nb) && (r < kp) && (k < Km);
  __shared__ double2 sg[MB][CH][16];
  double2 o1 = cx(0, 0), o2 = o1, o3 = o1, o4 = o1;
  const size_t base = static_cast<size_t>(G.poff[m]) * G.Jhp + r;
  const double *pc = Pa + base, *hc = Ha + base;
  // even rows pair P with X+ and H with Y-, odd rows P with X- and H with Y+
  const int xo = odd ? 4 : 0, yo = odd ? 8 : 12;
  for (int j0 = 0; j0 < G.Jh; j0 += CH) {
    const int cnt = min(CH, G.Jh - j0);
    __syncthreads();
    if (bl < nb) {
      // slots in sg: X+ 0..3, X- 4..7, Y+ 8..11, Y- 12..15, from the blocks (p, part) = (0,X) (0,Y) (1,X) (1,Y)
      const double2 *src = Gb + (static_cast<size_t>(b) * G.M1 + m) * 16 * G.Jh;
      for (int i = tx; i < 16 * cnt; i += 64) {
        const int blk = i / (4 * cnt), rem = i - blk * 4 * cnt, jj = rem >> 2, c = rem & 3;
        const int slot = (blk == 0 ? 0 : blk == 1 ? 12 : blk == 2 ? 4 : 8) + c;
        sg[ty][jj][slot] = src[(static_cast<size_t>(blk) * G.Jh + j0 + jj) * 4 + c];
      }
    }
    __syncthreads();
    if (valid) {
      for (int jj = 0; jj < cnt; jj++) {
        const double P = pc[static_cast<size_t>(j0 + jj) * kp], H = hc[static_cast<size_t>(j0 + jj) * kp];
        const double2 *gx = sg[ty][jj];
        cfma(o1, gx[xo + 0], P);
        cfma(o1, gx[yo + 0], H);
        cfma(o2, gx[xo + 1], P);
        cfma(o2, gx[yo + 1], H);
        cfma(o3, gx[xo + 2], P);
        cfma(o4, gx[xo + 3], P);
        cfma(o4, gx[yo + 3], H);
      }
    }
  }
  if (!valid) return;
  anal_epilogue(G, mode, stage, dt, nu4, b, bl, m + k, q + k, o1, o2, o3, o4, x0, xs, acc, raw, 0);
}

// Inverse FFTs of the four Fourier fields to the grid: diagnostics (h = phi/g, u = U/cos, v = V/cos, zeta;
// sphere.cpp:486-508) or the raw rows U, V, zeta, phi.
template <int R1, int R2>
__global__ __launch_bounds__(256) void k_to_grid(Geo G, int mode, const double2 *__restrict__ Fs, const double2 *__restrict__ twg,
                                                 const double *__restrict__ mu, double *__restrict__ out, int b0) {
  using F = Fft2<R1, R2>;
  constexpr int L = F::L, P = F::P;
  extern __shared__ double2 lds[];
  const int K = 2 * L, M1 = G.M1, tid = threadIdx.x;
  const int jh = blockIdx.x, b = b0 + blockIdx.y;
  const int jn = G.Jh + jh, js = G.Jh - 1 - jh;
  double2 *buf = lds, *tw = lds + 10 * P;
  for (int i = tid; i < K; i += blockDim.x) tw[i] = twg[i];
  __syncthreads();
  for (int idx = tid; idx < 8 * L; idx += blockDim.x) {
    const int r = idx / L, mp = idx - r * L;
    const int j = (r & 1) ? js : jn;
    const double2 *src = Fs + (static_cast<size_t>(b) * G.J + j) * M1 * 4 + (r >> 1);
    buf[r * P + F::phys(mp)] = irfft_entry(src, mp, L, G.N, tw);
  }
  __syncthreads();
  fft2<R1, R2, +1>(buf, 8, tw);
  const double c = sqrt(1 - mu[jn] * mu[jn]);
  const size_t plane = static_cast<size_t>(G.J) * K;
  double *o = out + static_cast<size_t>(b) * 4 * plane;
  for (int idx = tid; idx < 2 * L; idx += blockDim.x) {
    const int s = idx / L, k = idx - s * L;
    const int j = s ? js : jn, pk = F::phys(F::pos(k));
    const double2 U = buf[(0 + s) * P + pk], V = buf[(2 + s) * P + pk], Z = buf[(4 + s) * P + pk], Ph = buf[(6 + s) * P + pk];
    const size_t i = static_cast<size_t>(j) * K + 2 * k;
    if (mode == GRID_DIAG) {
      o[i] = Ph.x / G.g;
      o[i + 1] = Ph.y / G.g;
      o[plane + i] = U.x / c;
      o[plane + i + 1] = U.y / c;
      o[2 * plane + i] = V.x / c;
      o[2 * plane + i + 1] = V.y / c;
      o[3 * plane + i] = Z.x;
      o[3 * plane + i + 1] = Z.y;
    } else {
      o[i] = U.x;
      o[i + 1] = U.y;
      o[plane + i] = V.x;
      o[plane + i + 1] = V.y;
      o[2 * plane + i] = Z.x;
      o[2 * plane + i + 1] = Z.y;
      o[3 * plane + i] = Ph.x;
      o[3 * plane + i + 1] = Ph.y;
    }
  }
}

namespace {
// Neumaier-compensated accumulator
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
}  // namespace

// Area means and extremes (sphere.cpp:438-447,514-590) in two launches.  k_reduce_rows: one wavefront per latitude
// row (four rows per block), lanes along the longitudes with coalesced loads, each lane's Neumaier sums combined over
// the wavefront in lane order; rowred[b][q][j] for the 9 quantities q = kinetic, potential, potential enstrophy,
// max wind, max h, err num, err den, max |h - ref|, max |ref|.  k_reduce_combine: the rows of one member in order,
// as the reference sums them, into red[b][0..8].
__global__ __launch_bounds__(256) void k_reduce_rows(Geo G, const double *__restrict__ grid, const double *__restrict__ fgrid,
                                                     const double *__restrict__ hs, const double *__restrict__ ref_phi,
                                                     double *__restrict__ rowred, int b0, const int *__restrict__ mflags) {
  const int b = b0 + blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j = 4 * blockIdx.x + wave, K = G.K;
  __shared__ double part[4][9][65];
  if (j < G.J) {
    const size_t plane = static_cast<size_t>(G.J) * K;
    const double *hh = grid + static_cast<size_t>(b) * 4 * plane + static_cast<size_t>(j) * K;
    const double *uu = hh + plane, *vv = hh + 2 * plane, *zz = hh + 3 * plane;
    const double *ff = fgrid + static_cast<size_t>(b) * plane + static_cast<size_t>(j) * K;
    const double *ss = hs + static_cast<size_t>(mflags[b] >> 2) * plane + static_cast<size_t>(j) * K;
    const double *rr = ref_phi + static_cast<size_t>(b) * plane + static_cast<size_t>(j) * K;
    Kahan ke, pe, en, num, den;
    double wmax = 0, hmax = 0, mn = 0, mr = 0;
    const double g = G.g;
    for (int k = lane; k < K; k += 64) {
      const double h = hh[k], u = uu[k], v = vv[k], z = zz[k] + ff[k];
      ke.add(0.5 * h * (u * u + v * v));
      pe.add(0.5 * g * h * h + g * h * ss[k]);
      en.add(0.5 * z * z / h);
      wmax = fmax(wmax, hypot(u, v));
      hmax = fmax(hmax, h);
      const double rf = rr[k] / g, d = h - rf;
      num.add(d * d);
      den.add(rf * rf);
      mn = fmax(mn, fabs(d));
      mr = fmax(mr, fabs(rf));
    }
    const double v9[9] = {ke.value(), pe.value(), en.value(), wmax, hmax, num.value(), den.value(), mn, mr};
#pragma unroll
    for (int q = 0; q < 9; q++) part[wave][q][lane] = v9[q];
  }
  __syncthreads();
  if (j < G.J && lane < 9) {
    const int q = lane;
    double out = 0;
    if (q == 3 || q == 4 || q == 7 || q == 8) {
      for (int l = 0; l < 64; l++) out = fmax(out, part[wave][q][l]);
    } else {
      Kahan s;
      for (int l = 0; l < 64; l++) s.add(part[wave][q][l]);
      out = s.value();
    }
    rowred[(static_cast<size_t>(b) * 9 + q) * G.J + j] = out;
  }
}

__global__ __launch_bounds__(64) void k_reduce_combine(Geo G, const double *__restrict__ rowred, const double *__restrict__ w,
                                                       double *__restrict__ red, int b0) {
  const int b = b0 + blockIdx.x, q = threadIdx.x, K = G.K;
  if (q >= 9) return;
  // 0..2 area means (x 1/2, sphere.cpp:520-533), 3, 4 maxima, 5, 6 weighted sums (sphere.cpp:579-580), 7, 8 maxima
  const double *rows = rowred + (static_cast<size_t>(b) * 9 + q) * G.J;
  double out;
  if (q < 3 || q == 5 || q == 6) {
    Kahan s;
    for (int r = 0; r < G.J; r++) s.add(q < 3 ? w[r] * rows[r] / K : w[r] * rows[r]);
    out = q < 3 ? 0.5 * s.value() : s.value();
  } else {
    out = 0;
    for (int r = 0; r < G.J; r++) out = fmax(out, rows[r]);
  }
  red[static_cast<size_t>(b) * 16 + q] = out;
}

// Sum of |coefficient| over zeta, delta, phi of one member (the finiteness check of sphere.cpp:449-455).
__global__ __launch_bounds__(256) void k_abs_sum(const double2 *__restrict__ x0, size_t count, double *__restrict__ out) {
  __shared__ double part[256];
  double s = 0;
  for (size_t i = threadIdx.x; i < count; i += blockDim.x) s += hypot(x0[i].x, x0[i].y);
  part[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) part[threadIdx.x] += part[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0];
}

// Tracer grid products (sphere.cpp:255-258): in3 = U q, V q, q delta from raw grids U, V, zeta, q and the delta grid.
__global__ void k_tracer_products(const double *__restrict__ g4, const double *__restrict__ dgrid, double *__restrict__ in3, size_t plane) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= plane) return;
  const double U = g4[i], V = g4[plane + i], q = g4[3 * plane + i], d = dgrid[i];
  in3[i] = U * q;
  in3[plane + i] = V * q;
  in3[2 * plane + i] = q * d;
}

__device__ inline int degree_of(const Geo &G, int i, int &m) {
  int mm = 0;
  while (mm + 1 < G.M1 && G.moff[mm + 1] <= i) mm++;
  m = mm;
  return i - G.moff[mm] + mm;
}

// RK4 stage update of the tracer with dq = corr - flux (sphere.cpp:261, :423-433) and the del^4 filter of q at the
// last stage (:392-405); raw = the analysis outputs (o1 at [0, S), o4 at [3S, 4S)).
__global__ void k_tracer_rk4(Geo G, int stage, double dt, double nu4, const double2 *__restrict__ raw, double2 *__restrict__ q0,
                             double2 *__restrict__ qs, double2 *__restrict__ qacc) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.S) return;
  const double2 k = csub(raw[3 * static_cast<size_t>(G.S) + i], raw[i]);
  const double2 s0 = q0[i];
  if (stage == 1) {
    qacc[i] = k;
    qs[i] = cadd(s0, cscale(k, 0.5 * dt));
  } else if (stage == 2) {
    qacc[i] = cadd(qacc[i], cscale(k, 2.0));
    qs[i] = cadd(s0, cscale(k, 0.5 * dt));
  } else if (stage == 3) {
    qacc[i] = cadd(qacc[i], cscale(k, 2.0));
    qs[i] = cadd(s0, cscale(k, dt));
  } else {
    double2 nx = cadd(s0, cscale(cadd(qacc[i], k), dt / 6));
    int m;
    const int n = degree_of(G, i, m);
    if (nu4 > 0 && n >= 1) {
      const double kq = n * (n + 1.0) / (G.a * G.a);
      nx = cscale(nx, exp(-nu4 * kq * kq * dt));
    }
    q0[i] = nx;
    qs[i] = nx;
  }
}

// del^4 filter of zeta, delta, phi of member 0 (sphere.cpp:392-405) for advect-only steps, whose frozen dynamics
// skip the stage kernels that would apply it
__global__ void k_state_filter(Geo G, double dt, double nu4, double2 *__restrict__ x0, double2 *__restrict__ xs) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.S) return;
  int m;
  const int n = degree_of(G, i, m);
  if (n < 1) return;
  const double a2 = G.a * G.a, l = n * (n + 1.0);
  const double kv = (l - 2) / a2, kq = l / a2;
  const double fv = exp(-nu4 * kv * kv * dt), fp = exp(-nu4 * kq * kq * dt);
  for (int fld = 0; fld < 3; fld++) {
    const size_t j = static_cast<size_t>(fld) * G.S + i;
    const double2 v = cscale(x0[j], fld == 2 ? fp : fv);
    x0[j] = v;
    xs[j] = v;
  }
}

// sum of |coefficient| of member b of a strided state array (check_finite)
__global__ __launch_bounds__(256) void k_abs_sum_member(Geo G, int b, const double2 *__restrict__ x, double *__restrict__ out) {
  __shared__ double part[256];
  double s = 0;
  const size_t n = 3 * static_cast<size_t>(G.S);
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double2 v = x[G.at(b, static_cast<int>(i / G.S), i % G.S)];
    s += hypot(v.x, v.y);
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) part[threadIdx.x] += part[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0];
}

// member b, field f of a strided state array <-> a contiguous [S] buffer
__global__ void k_member_field(Geo G, int b, int f, double2 *__restrict__ x0, double2 *__restrict__ xs, double2 *__restrict__ buf, int to_state) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= G.S) return;
  const size_t i = G.at(b, f, q);
  if (to_state) {
    x0[i] = buf[q];
    xs[i] = buf[q];
  } else {
    buf[q] = x0[i];
  }
}

// dst (layout Gd) <- src (layout Gs), every member, field and coefficient
__global__ void k_relayout(Geo Gs, Geo Gd, const double2 *__restrict__ src, double2 *__restrict__ dst) {
  const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t n3 = 3 * static_cast<size_t>(Gs.S);
  if (e >= n3 * Gs.B) return;
  const int b = static_cast<int>(e / n3), f = static_cast<int>((e % n3) / Gs.S);
  const size_t q = e % Gs.S;
  dst[Gd.at(b, f, q)] = src[Gs.at(b, f, q)];
}

// ------------------------------------------------------------------------------------------- host side

static void tracer_grid_device(SphereCore &c, double *dst);

SphereCore::SphereCore(int truncation, const lmw_sphere_params &p, int members, int prec) : prm(p), precision(LMW_FP64) {
  if (truncation < 4 || truncation > 170) invalid("truncation must be between 4 and 170");
  if (members < 1) invalid("an ensemble needs at least one member");
  N = truncation;
  K = 4;
  while (K < 3 * N + 1) K *= 2;
  J = K / 2;
  logK = 0;
  while ((1 << logK) < K) logK++;
  Jh = J / 2;
  Jhp = (Jh + 15) / 16 * 16;
  M1 = N + 1;
  S = M1 * (M1 + 1) / 2;
  B = members;
  h_moff.resize(M1);
  h_poff.resize(M1);
  h_kpe.resize(M1);
  h_kpo.resize(M1);
  R = 0;
  for (int m = 0; m <= N; m++) {
    const int Km = N - m + 1;
    h_moff[m] = m * M1 - m * (m - 1) / 2;
    h_kpe[m] = ((Km + 1) / 2 + 15) / 16 * 16;
    h_kpo[m] = (Km / 2 + 15) / 16 * 16;
    h_poff[m] = R;
    R += h_kpe[m] + h_kpo[m];
  }
  Tmax = h_kpe[0] + h_kpo[0];
  LMW_HIP(hipStreamCreate(&stream));
  moff.alloc(M1);
  poff.alloc(M1);
  kpe.alloc(M1);
  kpo.alloc(M1);
  moff.upload(h_moff.data(), M1, 0, stream);
  poff.upload(h_poff.data(), M1, 0, stream);
  kpe.upload(h_kpe.data(), M1, 0, stream);
  kpo.upload(h_kpo.data(), M1, 0, stream);
  mu.alloc(J);
  w.alloc(J);
  const size_t tab = static_cast<size_t>(R) * Jhp;
  Ps.alloc(tab);
  Hs.alloc(tab);
  Pa.alloc(tab);
  Ha.alloc(tab);
  std::vector<double2> htw(K);
  for (int q = 0; q < K; q++) {
    const long double ang = 2.0L * 3.141592653589793238462643383279502884L * q / K;
    htw[q] = make_double2(static_cast<double>(cosl(ang)), static_cast<double>(-sinl(ang)));
  }
  tw.alloc(K);
  tw.upload(htw.data(), K, 0, stream);
  std::vector<float2> htw32(K);  // the same twiddles rounded once to FP32, for the mixed-precision grid kernel
  for (int q = 0; q < K; q++) htw32[q] = make_float2(static_cast<float>(htw[q].x), static_cast<float>(htw[q].y));
  tw32.alloc(K);
  tw32.upload(htw32.data(), K, 0, stream);
  std::vector<double> hkn(N + 2);  // the inverse Laplacian of the synthesis kernels (sphere.cpp:185-190), same expression
  for (int n = 0; n <= N + 1; n++) hkn[n] = n >= 1 ? -prm.radius * prm.radius / (n * (n + 1.0)) : 0.0;
  knt.alloc(N + 2);
  knt.upload(hkn.data(), N + 2, 0, stream);
  sync();
  const size_t nstate = static_cast<size_t>(B) * 3 * S;
  x0.alloc(nstate);
  xs.alloc(nstate);
  acc.alloc(nstate);
  Fs.alloc(static_cast<size_t>(B) * J * M1 * 4);
  Gb.alloc(static_cast<size_t>(B) * M1 * Jh * 16);
  const size_t plane = static_cast<size_t>(J) * K;
  f.alloc(B * plane);
  hs.alloc(B * plane);
  ref_phi.alloc(B * plane);
  grid.alloc(B * 4 * plane);
  in3.alloc(3 * plane);
  spec_io.alloc(4 * static_cast<size_t>(S));
  red.alloc(static_cast<size_t>(B) * 16);
  rowred.alloc(static_cast<size_t>(B) * 9 * J);
  t.assign(B, 0.0);
  tilt.assign(B, 0.0);
  has_orography.assign(B, 0);
  hs_slot.resize(B);
  for (int b = 0; b < B; b++) hs_slot[b] = b;
  hs_refs.assign(B, 1);
  hs_kind.assign(B, 0);
  hs_key.assign(B, 0);
  mflags.alloc(B);
  Geo G = geo();
  hipLaunchKernelGGL(k_gauss, dim3(1), dim3(256), 0, stream, J, mu.p, w.p);
  LMW_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_tables, dim3(M1), dim3(Jhp), 0, stream, G, mu.p, Ps.p, Hs.p, Pa.p, Ha.p);
  LMW_LAUNCH_CHECK();
  for (int b = 0; b < B; b++) set_axis_tilt(b, 0.0);
  set_precision(prec);
  // FP64 Legendre kernels of groups of two or more full 16-member tiles: the FP64 MFMA kernels (row chunk 32, 4
  // waves), which the autotuner picks at T42, T85 and T170 with 32 to 128 members; smaller groups and single members
  // keep the FMA kernels with 4 members per block
  if (B >= 32) {
    tile_fp64_synth = kFp64Tiles;
    tile_fp64_anal = kFp64Tiles;
    build_fp64_mfma_tables();
  }
  sync();
}

SphereCore::~SphereCore() {
  if (graph) (void)hipGraphExecDestroy(graph);
  if (graph_many) (void)hipGraphExecDestroy(graph_many);
  for (int p = 1; p < kMaxParts; p++) {
    if (part_stream[p]) (void)hipStreamDestroy(part_stream[p]);
    if (part_join[p]) (void)hipEventDestroy(part_join[p]);
  }
  if (part_fork) (void)hipEventDestroy(part_fork);
  if (stream) (void)hipStreamDestroy(stream);
}

Geo SphereCore::geo() const {
  Geo G;
  G.N = N;
  G.M1 = M1;
  G.J = J;
  G.K = K;
  G.logK = logK;
  G.Jh = Jh;
  G.Jhp = Jhp;
  G.B = B;
  G.S = S;
  G.Tmax = Tmax;
  if (member_minor) {
    G.sb = 1;
    G.sf = static_cast<long long>(S) * B;
    G.sq = B;
  } else {
    G.sb = 3LL * S;
    G.sf = S;
    G.sq = 1;
  }
  G.a = prm.radius;
  G.g = prm.gravity;
  G.omega = prm.omega;
  G.moff = moff.p;
  G.poff = poff.p;
  G.kpe = kpe.p;
  G.kpo = kpo.p;
  return G;
}

void SphereCore::sync() { LMW_HIP(hipStreamSynchronize(stream)); }

void SphereCore::update_flags(int b) {
  const int fl = (tilt[b] == 0.0 ? 1 : 0) | (has_orography[b] ? 0 : 2) | (hs_slot[b] << 2);
  mflags.upload(&fl, 1, b, stream);
  sync();  // the source is a stack variable
}

void SphereCore::set_axis_tilt(int b, double alpha) {
  tilt[b] = alpha;
  update_flags(b);
  const size_t plane = static_cast<size_t>(J) * K;
  hipLaunchKernelGGL(k_coriolis, dim3(J), dim3(256), 0, stream, geo(), mu.p, f.p + b * plane, alpha);
  LMW_LAUNCH_CHECK();
}

void SphereCore::point_orography(int b, int p) {
  hs_refs[hs_slot[b]]--;
  hs_slot[b] = p;
  hs_refs[p]++;
}

// A plane that member b alone refers to, its contents unknown.  One exists: there are B planes and the other B - 1
// members refer to at most B - 1 of them.
int SphereCore::own_orography(int b) {
  int p = hs_slot[b];
  if (hs_refs[p] != 1) {
    for (p = 0; p < B && hs_refs[p] != 0; p++) {
    }
    if (p == B) runtime("no free orography plane (the plane reference counts are inconsistent)");
    point_orography(b, p);
  }
  hs_kind[p] = 0;
  return p;
}

// set_orography_grid (sphere.cpp:105-109): member b shares a plane that already holds the same bits (checked on the
// device copy), else gets one of its own
void SphereCore::set_orography(int b, const double *hsh) { set_orography_run(b, 1, hsh, nullptr); }

void SphereCore::set_orography_run(int b0, int nb, const double *hsh, const double *dev_src) {
  const size_t plane = static_cast<size_t>(J) * K;
  uint64_t key = 1469598103934665603ull;  // FNV-1a over the 64-bit words
  for (size_t i = 0; i < plane; i++) {
    uint64_t bits;
    std::memcpy(&bits, hsh + i, sizeof bits);
    key = (key ^ bits) * 1099511628211ull;
  }
  int found = -1;
  std::vector<double> tmp;
  for (int p = 0; p < B && found < 0; p++) {
    if (hs_kind[p] != 1 || hs_key[p] != key) continue;
    tmp.resize(plane);
    hs.download(tmp.data(), plane, static_cast<size_t>(p) * plane, stream);
    sync();
    if (std::memcmp(tmp.data(), hsh, plane * sizeof(double)) == 0) found = p;
  }
  if (found < 0) {
    found = own_orography(b0);
    if (dev_src)
      LMW_HIP(hipMemcpyAsync(hs.p + static_cast<size_t>(found) * plane, dev_src, plane * sizeof(double), hipMemcpyDeviceToDevice, stream));
    else
      hs.upload(hsh, plane, static_cast<size_t>(found) * plane, stream);
    hs_kind[found] = 1;
    hs_key[found] = key;
  }
  for (int b = b0; b < b0 + nb; b++) {
    point_orography(b, found);
    has_orography[b] = 1;
    update_flags(b);
  }
}

template <int MB>
static void synth_launch(SphereCore &c, const double2 *spec, int b0, int nb, bool scratch) {
  Geo G = c.geo();
  if (scratch) {  // member-major scratch buffers
    G.sb = 3LL * c.S;
    G.sf = c.S;
    G.sq = 1;
  }
  dim3 grd(c.M1, (c.Jh + 63) / 64, (nb + MB - 1) / MB), blk(64, MB);
  hipLaunchKernelGGL(k_synth<MB>, grd, blk, 0, c.stream, G, c.Ps.p, c.Hs.p, spec, c.Fs.p, b0, nb);
  LMW_LAUNCH_CHECK();
}

void SphereCore::launch_synth(const double2 *spec, int b0, int nb, bool scratch) {
  if (tile_fp64_synth >= kFp64Tiles) {  // the FP64 MFMA kernel
    launch_synth_fp64_mfma(tile_fp64_synth == kFp64Tiles ? 32 : 64, spec, b0, nb, scratch);
    return;
  }
  switch (kFp64MB[tile_fp64_synth]) {
    case 2: synth_launch<2>(*this, spec, b0, nb, scratch); break;
    case 8: synth_launch<8>(*this, spec, b0, nb, scratch); break;
    default: synth_launch<4>(*this, spec, b0, nb, scratch); break;
  }
}

// LDS of the grid kernels: 10 padded rows of K/2 plus the K twiddles
static size_t grid_lds_bytes(int K) {
  const int L = K / 2, R2 = L >= 128 ? (L == 256 ? 16 : 8) : L >= 32 ? (L == 64 ? 8 : 4) : L == 16 ? 4 : 2;
  return (10 * static_cast<size_t>(L + L / R2) + K) * sizeof(double2);
}

// L = K/2 -> (R1, R2): 256 (16,16), 128 (16,8), 64 (8,8), 32 (8,4), 16 (4,4), 8 (4,2)
#define LMW_DISPATCH_L(L, KERNEL, ...)                                               \
  switch (L) {                                                                       \
    case 256: hipLaunchKernelGGL((KERNEL<16, 16>), __VA_ARGS__); break;             \
    case 128: hipLaunchKernelGGL((KERNEL<16, 8>), __VA_ARGS__); break;              \
    case 64: hipLaunchKernelGGL((KERNEL<8, 8>), __VA_ARGS__); break;                \
    case 32: hipLaunchKernelGGL((KERNEL<8, 4>), __VA_ARGS__); break;                \
    case 16: hipLaunchKernelGGL((KERNEL<4, 4>), __VA_ARGS__); break;                \
    case 8: hipLaunchKernelGGL((KERNEL<4, 2>), __VA_ARGS__); break;                 \
    default: invalid("unsupported grid size");                                     \
  }
// the same with a storage type as third template argument
#define LMW_DISPATCH_LT(L, KERNEL, TS, ...)                                          \
  switch (L) {                                                                       \
    case 256: hipLaunchKernelGGL((KERNEL<16, 16, TS>), __VA_ARGS__); break;         \
    case 128: hipLaunchKernelGGL((KERNEL<16, 8, TS>), __VA_ARGS__); break;          \
    case 64: hipLaunchKernelGGL((KERNEL<8, 8, TS>), __VA_ARGS__); break;            \
    case 32: hipLaunchKernelGGL((KERNEL<8, 4, TS>), __VA_ARGS__); break;            \
    case 16: hipLaunchKernelGGL((KERNEL<4, 4, TS>), __VA_ARGS__); break;            \
    case 8: hipLaunchKernelGGL((KERNEL<4, 2, TS>), __VA_ARGS__); break;             \
    default: invalid("unsupported grid size");                                     \
  }

void SphereCore::launch_tend_grid(int b0, int nb, int mixed) {
  if (mixed) {  // FP32 rows, FP32 arithmetic: half the LDS of the FP64 kernel
    LMW_DISPATCH_LT(K / 2, k_tend_grid, float2, dim3(Jh, nb), dim3(256), grid_lds_bytes(K) / 2, stream, geo(), Fs32.p, f.p, hs.p, tw32.p,
                    mu.p, w.p, Gb32.p, b0, mflags.p)
  } else {
    LMW_DISPATCH_LT(K / 2, k_tend_grid, double2, dim3(Jh, nb), dim3(256), grid_lds_bytes(K), stream, geo(), Fs.p, f.p, hs.p, tw.p, mu.p, w.p,
                    Gb.p, b0, mflags.p)
  }
  LMW_LAUNCH_CHECK();
}

template <int MB>
static void anal_launch(SphereCore &c, int mode, int stage, double dt, int b0, int nb, double2 *raw_out) {
  dim3 grd(c.M1, (c.Tmax + 63) / 64, (nb + MB - 1) / MB), blk(64, MB);
  hipLaunchKernelGGL(k_anal<MB>, grd, blk, 0, c.stream, c.geo(), c.Pa.p, c.Ha.p, c.Gb.p, mode, stage, dt, c.prm.diffusion, c.x0.p, c.xs.p, c.acc.p,
                     raw_out, b0, nb);
  LMW_LAUNCH_CHECK();
}

void SphereCore::launch_anal(int mode, int stage, double dt, int b0, int nb, double2 *raw_out) {
  if (tile_fp64_anal >= kFp64Tiles) {  // the FP64 MFMA kernel
    launch_anal_fp64_mfma(tile_fp64_anal == kFp64Tiles ? 4 : 8, mode, stage, dt, b0, nb, raw_out);
    return;
  }
  switch (kFp64MB[tile_fp64_anal]) {
    case 2: anal_launch<2>(*this, mode, stage, dt, b0, nb, raw_out); break;
    case 8: anal_launch<8>(*this, mode, stage, dt, b0, nb, raw_out); break;
    default: anal_launch<4>(*this, mode, stage, dt, b0, nb, raw_out); break;
  }
}

void SphereCore::launch_prep(int mode, int b0, int nb) {
  // in3 holds the input grids of members b0.. at offset 0 (k_prep indexes it by the local member)
  LMW_DISPATCH_L(K / 2, k_prep, dim3(Jh, nb), dim3(256), grid_lds_bytes(K), stream, geo(), mode, in3.p, tw.p, mu.p, w.p, Gb.p, b0);
  LMW_LAUNCH_CHECK();
}

void SphereCore::launch_to_grid(int mode, double *out, int b0, int nb) {
  LMW_DISPATCH_L(K / 2, k_to_grid, dim3(Jh, nb), dim3(256), grid_lds_bytes(K), stream, geo(), mode, Fs.p, tw.p, mu.p, out, b0);
  LMW_LAUNCH_CHECK();
}

void SphereCore::set_precision(int p) {
  if (p != LMW_FP64 && p != LMW_FP32 && p != LMW_BF16X3 && p != LMW_BF16X6)
    invalid("precision must be LMW_FP64, LMW_FP32, LMW_BF16X3 or LMW_BF16X6");
  if (p != LMW_FP64) build_mixed_tables();
  const bool minor = B > 1;  // groups of several members: member-minor in every precision (set at construction)
  if (minor != member_minor) {
    const Geo Gs = geo();
    member_minor = minor;
    const Geo Gd = geo();
    const size_t n = x0.n;
    DevArray<double2> tmp;
    tmp.alloc(n);
    for (DevArray<double2> *a : {&x0, &xs, &acc}) {
      hipLaunchKernelGGL(k_relayout, dim3((n + 255) / 256), dim3(256), 0, stream, Gs, Gd, a->p, tmp.p);
      LMW_LAUNCH_CHECK();
      LMW_HIP(hipMemcpyAsync(a->p, tmp.p, n * sizeof(double2), hipMemcpyDeviceToDevice, stream));
    }
    sync();
  }
  precision = p;
  drop_graph();
}

void SphereCore::set_state_grid(int b, const double *h, const double *u, const double *v) {
  const size_t plane = static_cast<size_t>(J) * K;
  in3.upload(h, plane, 0, stream);
  in3.upload(u, plane, plane, stream);
  in3.upload(v, plane, 2 * plane, stream);
  state_from_in3(b, 1);
}

// in3 large enough for the input grids of nb members
void SphereCore::ensure_in3(int nb) {
  const size_t need = static_cast<size_t>(nb) * 3 * J * K;
  if (in3.n < need) {
    sync();  // in3 may still be read by queued work
    in3.alloc(need);
  }
}

// The grids h, u, v of members b0 .. b0 + nb - 1 in in3 ([local member][3][J][K]) become their spectral state:
// set_state_grid (sphere.cpp:217-234) for nb members in one launch of each kernel.
void SphereCore::state_from_in3(int b0, int nb) {
  launch_prep(PREP_STATE, b0, nb);
  launch_anal(ANAL_STATE, 0, 0.0, b0, nb, nullptr);
  for (int b = b0; b < b0 + nb; b++) t[b] = 0;
  mark_reference_range(b0, nb);
}

// set_state_grid of members b0 .. b0 + nb - 1 from host grids [nb][J][K], kSetupChunk members per launch
void SphereCore::set_state_grids(int b0, int nb, const double *h, const double *u, const double *v) {
  const size_t plane = static_cast<size_t>(J) * K, pb = plane * sizeof(double);
  for (int c0 = 0; c0 < nb; c0 += kSetupChunk) {
    const int n = std::min(kSetupChunk, nb - c0);
    ensure_in3(n);
    const double *src[3] = {h, u, v};
    for (int f = 0; f < 3; f++)
      LMW_HIP(hipMemcpy2DAsync(in3.p + f * plane, 3 * pb, src[f] + static_cast<size_t>(c0) * plane, pb, pb, n, hipMemcpyHostToDevice, stream));
    state_from_in3(b0 + c0, n);
    sync();  // the next chunk reuses in3
  }
}

// The analytic set-ups of Williamson et al. (1992) (cpp/main.cpp:174-291) for members b0 .. b0 + nb - 1, evaluated
// and analysed on the device.  Test case 1 also sets the (single-member) tracer and the advect-only mode.
void SphereCore::setup(int b0, int nb, int which, double alpha) {
  if (which != LMW_SETUP_WILLIAMSON1 && which != LMW_SETUP_WILLIAMSON2 && which != LMW_SETUP_WILLIAMSON5 && which != LMW_SETUP_WILLIAMSON6)
    invalid("unknown set-up: LMW_SETUP_WILLIAMSON1, 2, 5 or 6");
  if (!std::isfinite(alpha)) invalid("angle must be a finite number");
  if (which == LMW_SETUP_WILLIAMSON1 && B != 1) invalid("a tracer needs a single-member sphere");
  const bool tilted = which == LMW_SETUP_WILLIAMSON1 || which == LMW_SETUP_WILLIAMSON2;
  const size_t plane = static_cast<size_t>(J) * K;
  for (int c0 = 0; c0 < nb; c0 += kSetupChunk) {
    const int n = std::min(kSetupChunk, nb - c0);
    double *cone = nullptr;  // the plane this launch writes test case 5's cone into, if no plane holds it yet
    if (which == LMW_SETUP_WILLIAMSON5) {
      int p = 0;
      while (p < B && !(hs_kind[p] == 2 && hs_refs[p] > 0)) p++;
      if (p == B) {
        p = own_orography(b0 + c0);
        hs_kind[p] = 2;
        cone = hs.p + static_cast<size_t>(p) * plane;
      }
      for (int b = b0 + c0; b < b0 + c0 + n; b++) {
        point_orography(b, p);
        has_orography[b] = 1;  // before the flags update of set_axis_tilt
      }
    }
    for (int b = b0 + c0; b < b0 + c0 + n; b++) set_axis_tilt(b, tilted ? alpha : 0.0);
    ensure_in3(n);
    hipLaunchKernelGGL(k_setup, dim3(J, n), dim3(256), 0, stream, geo(), which, alpha, mu.p, in3.p, cone);
    LMW_LAUNCH_CHECK();
    state_from_in3(b0 + c0, n);
    sync();
  }
  if (which == LMW_SETUP_WILLIAMSON1) {
    hipLaunchKernelGGL(k_setup, dim3(J, 1), dim3(256), 0, stream, geo(), SETUP_TRACER, alpha, mu.p, in3.p, nullptr);
    LMW_LAUNCH_CHECK();
    tracer_from_in3();
    set_advect_only(true);
    sync();
  }
}

void SphereCore::set_spectral_packed(int b, int field, const double2 *packed) {
  if (member_minor) {
    spec_io.upload(packed, S, 0, stream);
    hipLaunchKernelGGL(k_member_field, dim3((S + 255) / 256), dim3(256), 0, stream, geo(), b, field, x0.p, xs.p, spec_io.p, 1);
    LMW_LAUNCH_CHECK();
    sync();
    return;
  }
  const size_t off = (static_cast<size_t>(b) * 3 + field) * S;
  x0.upload(packed, S, off, stream);
  xs.upload(packed, S, off, stream);
}

void SphereCore::get_spectral_packed(int b, int field, double2 *packed) {
  if (member_minor) {
    hipLaunchKernelGGL(k_member_field, dim3((S + 255) / 256), dim3(256), 0, stream, geo(), b, field, x0.p, xs.p, spec_io.p, 0);
    LMW_LAUNCH_CHECK();
    spec_io.download(packed, S, 0, stream);
    sync();
    return;
  }
  x0.download(packed, S, (static_cast<size_t>(b) * 3 + field) * S, stream);
  sync();
}

void SphereCore::mark_reference(int b) { mark_reference_range(b, 1); }

// mark_reference (sphere.cpp:236-239) of members b0 .. b0 + nb - 1: the phi grids become the reference of error_l2
void SphereCore::mark_reference_range(int b0, int nb) {
  const size_t plane = static_cast<size_t>(J) * K, pb = plane * sizeof(double);
  launch_synth(x0.p, b0, nb);
  launch_to_grid(GRID_RAW, grid.p, b0, nb);  // rows of member b land at grid[b][0..3]
  LMW_HIP(hipMemcpy2DAsync(ref_phi.p + static_cast<size_t>(b0) * plane, pb, grid.p + (static_cast<size_t>(b0) * 4 + 3) * plane, 4 * pb, pb, nb,
                           hipMemcpyDeviceToDevice, stream));
  if (has_tracer) tracer_grid_device(*this, ref_q.p);  // sphere.cpp:238
}

void SphereCore::launch_stage(int stage, double dt, int b0, int nb) {
  if (precision == LMW_FP64) {
    launch_synth(xs.p, b0, nb);
    launch_tend_grid(b0, nb, 0);
    launch_anal(ANAL_TEND, stage, dt, b0, nb, nullptr);
  } else {  // only the Legendre contractions of the time step change precision
    launch_synth_mx(xs.p, b0, nb);
    launch_tend_grid(b0, nb, 1);
    launch_anal_mx(ANAL_TEND, stage, dt, b0, nb, nullptr);
  }
}

// Members are independent, so a group of several 16-member tiles is stepped as up to kMaxParts parts of whole tiles,
// each part's four stages a branch of its own: one part's kernels fill the gaps (tails, synchronisation phases) of
// another's (32 x T170: FP64 43.6 -> 51.3, BF16x6 124.5 -> 130.0 member-days/s).  Two parts beat four and eight for
// the larger groups of the uneven ensemble.  Single members and tracers keep one branch.
int SphereCore::step_parts() const {
  const int tiles = (B + 15) / 16;
  return has_tracer || advect_only ? 1 : std::min(tiles, kMaxParts);
}

void SphereCore::enqueue_step(double dt) { enqueue_steps(dt, 1); }

void SphereCore::enqueue_steps(double dt, int nsteps) {
  const int parts = step_parts();
  if (parts > 1) {
    const int tiles = (B + 15) / 16;
    if (!part_fork) {
      LMW_HIP(hipEventCreateWithFlags(&part_fork, hipEventDisableTiming));
      for (int p = 1; p < kMaxParts; p++) {
        LMW_HIP(hipStreamCreate(&part_stream[p]));
        LMW_HIP(hipEventCreateWithFlags(&part_join[p], hipEventDisableTiming));
      }
    }
    LMW_HIP(hipEventRecord(part_fork, stream));
    const hipStream_t main = stream;
    for (int p = parts - 1; p >= 0; p--) {
      const int b0 = 16 * (p * tiles / parts), b1 = std::min(B, 16 * ((p + 1) * tiles / parts));
      if (p > 0) {
        LMW_HIP(hipStreamWaitEvent(part_stream[p], part_fork, 0));
        stream = part_stream[p];  // the launch functions enqueue on `stream`
      }
      for (int i = 0; i < nsteps; i++)
        for (int stage = 1; stage <= 4; stage++) launch_stage(stage, dt, b0, b1 - b0);
      if (p > 0) {
        LMW_HIP(hipEventRecord(part_join[p], stream));
        stream = main;
      }
    }
    for (int p = 1; p < parts; p++) LMW_HIP(hipStreamWaitEvent(stream, part_join[p], 0));
    return;
  }
  for (int i = 0; i < nsteps; i++) {
    // the tracer tendency reads the stage state before the dynamics' stage kernel advances it (sphere.cpp:412-427)
    for (int stage = 1; stage <= 4; stage++) {
      if (has_tracer) launch_tracer_tendency();
      if (!advect_only) launch_stage(stage, dt, 0, B);
      if (has_tracer) launch_tracer_update(stage, dt);
    }
    if (advect_only && prm.diffusion > 0) launch_state_filter(dt);
  }
}

void SphereCore::drop_graph() {
  for (hipGraphExec_t *ge : {&graph, &graph_many})
    if (*ge) {
      LMW_HIP(hipGraphExecDestroy(*ge));
      *ge = nullptr;
    }
}

void SphereCore::launch_tracer_tendency() {
  const size_t plane = static_cast<size_t>(J) * K;
  LMW_HIP(hipMemcpyAsync(tq.p, xs.p, 2 * static_cast<size_t>(S) * sizeof(double2), hipMemcpyDeviceToDevice, stream));
  LMW_HIP(hipMemcpyAsync(tq.p + 2 * S, qs.p, S * sizeof(double2), hipMemcpyDeviceToDevice, stream));
  launch_synth(tq.p, 0, 1, true);
  launch_to_grid(GRID_RAW, grid.p, 0, 1);  // U, V, zeta, q
  LMW_HIP(hipMemcpyAsync(tq.p + 2 * S, xs.p + S, S * sizeof(double2), hipMemcpyDeviceToDevice, stream));
  launch_synth(tq.p, 0, 1, true);
  launch_to_grid(GRID_RAW, grid2.p, 0, 1);  // plane 3: delta
  hipLaunchKernelGGL(k_tracer_products, dim3((plane + 255) / 256), dim3(256), 0, stream, grid.p, grid2.p + 3 * plane, in3.p, plane);
  LMW_LAUNCH_CHECK();
  launch_prep(PREP_TRACER, 0, 1);
  launch_anal(ANAL_RAW, 0, 0.0, 0, 1, spec_io.p);
}

void SphereCore::launch_tracer_update(int stage, double dt) {
  hipLaunchKernelGGL(k_tracer_rk4, dim3((S + 255) / 256), dim3(256), 0, stream, geo(), stage, dt, prm.diffusion, spec_io.p, q0.p, qs.p, qacc.p);
  LMW_LAUNCH_CHECK();
}

void SphereCore::launch_state_filter(double dt) {
  hipLaunchKernelGGL(k_state_filter, dim3((S + 255) / 256), dim3(256), 0, stream, geo(), dt, prm.diffusion, x0.p, xs.p);
  LMW_LAUNCH_CHECK();
}

static void tracer_buffers(SphereCore &c) {
  if (c.q0.n) return;
  const size_t plane = static_cast<size_t>(c.J) * c.K;
  c.q0.alloc(c.S);
  c.qs.alloc(c.S);
  c.qacc.alloc(c.S);
  c.tq.alloc(3 * static_cast<size_t>(c.S));
  c.grid2.alloc(4 * plane);
  c.ref_q.alloc(plane);
}

// q of member 0 synthesized to the grid at dst (device)
static void tracer_grid_device(SphereCore &c, double *dst) {
  const size_t plane = static_cast<size_t>(c.J) * c.K;
  LMW_HIP(hipMemsetAsync(c.tq.p, 0, 2 * static_cast<size_t>(c.S) * sizeof(double2), c.stream));
  LMW_HIP(hipMemcpyAsync(c.tq.p + 2 * c.S, c.q0.p, c.S * sizeof(double2), hipMemcpyDeviceToDevice, c.stream));
  c.launch_synth(c.tq.p, 0, 1, true);
  c.launch_to_grid(GRID_RAW, c.grid2.p, 0, 1);
  LMW_HIP(hipMemcpyAsync(dst, c.grid2.p + 3 * plane, plane * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
}

void SphereCore::set_tracer_spectral(const double2 *packed) {
  if (B != 1) invalid("a tracer needs a single-member sphere");
  tracer_buffers(*this);
  q0.upload(packed, S, 0, stream);
  qs.upload(packed, S, 0, stream);
  has_tracer = true;
  drop_graph();
  sync();
}

void SphereCore::set_tracer_grid(const double *q) {
  if (B != 1) invalid("a tracer needs a single-member sphere");
  const size_t plane = static_cast<size_t>(J) * K;
  in3.upload(q, plane, 0, stream);
  tracer_from_in3();
}

// the tracer grid in plane 0 of in3 becomes the tracer's spectral state and its reference (sphere.cpp:241-249)
void SphereCore::tracer_from_in3() {
  tracer_buffers(*this);
  launch_prep(PREP_SCALAR, 0, 1);
  launch_anal(ANAL_RAW, 0, 0.0, 0, 1, spec_io.p);
  LMW_HIP(hipMemcpyAsync(q0.p, spec_io.p + 3 * static_cast<size_t>(S), S * sizeof(double2), hipMemcpyDeviceToDevice, stream));
  LMW_HIP(hipMemcpyAsync(qs.p, q0.p, S * sizeof(double2), hipMemcpyDeviceToDevice, stream));
  has_tracer = true;
  drop_graph();
  tracer_grid_device(*this, ref_q.p);  // sphere.cpp:245
  sync();
}

void SphereCore::set_advect_only(bool on) {
  if (on != advect_only) drop_graph();
  advect_only = on;
}

void SphereCore::tracer_grid(double *out) {
  if (!has_tracer) runtime("no tracer: set one with tracer() first");
  const size_t plane = static_cast<size_t>(J) * K;
  tracer_grid_device(*this, grid2.p);
  grid2.download(out, plane, 0, stream);
  sync();
}

double SphereCore::tracer_abs_sum() {
  hipLaunchKernelGGL(k_abs_sum, dim3(1), dim3(256), 0, stream, q0.p, static_cast<size_t>(S), red.p + 14);
  LMW_LAUNCH_CHECK();
  double s;
  red.download(&s, 1, 14, stream);
  sync();
  return s;
}

void SphereCore::steps(double dt, int64_t count) {
  for (int64_t i = 0; i < count; i += kGraphSteps) enqueue_steps(dt, static_cast<int>(std::min<int64_t>(kGraphSteps, count - i)));
}

// Replays of a graph of kGraphSteps steps, then of the one-step graph for the rest.  A graph of several steps forks
// the member parts once, so each part runs its steps without waiting for the others at every step, and the device
// crosses fewer graph boundaries (32 x T170 BF16x6: 130.2 member-days/s with one step per graph, 136.2 with 4, 136.9
// with 8).
void SphereCore::steps_graph(double dt, int64_t count) {
  if (count <= 0) return;
  auto capture = [&](hipGraphExec_t &ge, int nsteps) {
    hipGraph_t g;
    LMW_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    enqueue_steps(dt, nsteps);
    LMW_HIP(hipStreamEndCapture(stream, &g));
    LMW_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    LMW_HIP(hipGraphDestroy(g));
  };
  if (graph_dt != dt) drop_graph();
  graph_dt = dt;
  int64_t i = 0;
  if (count >= kGraphSteps) {
    if (!graph_many) capture(graph_many, kGraphSteps);
    for (; i + kGraphSteps <= count; i += kGraphSteps) LMW_HIP(hipGraphLaunch(graph_many, stream));
  }
  if (i < count && !graph) capture(graph, 1);
  for (; i < count; i++) LMW_HIP(hipGraphLaunch(graph, stream));
}

void SphereCore::steps_timed(double dt, int64_t count, double ms[3]) {
  if (count <= 0) return;
  std::vector<hipEvent_t> ev(static_cast<size_t>(count) * 16);
  for (auto &e : ev) LMW_HIP(hipEventCreate(&e));
  size_t q = 0;
  for (int64_t i = 0; i < count; i++)
    for (int stage = 1; stage <= 4; stage++, q += 4) {
      const bool mx = precision != LMW_FP64;
      LMW_HIP(hipEventRecord(ev[q], stream));
      if (mx) launch_synth_mx(xs.p, 0, B);
      else launch_synth(xs.p, 0, B);
      LMW_HIP(hipEventRecord(ev[q + 1], stream));
      launch_tend_grid(0, B, mx ? 1 : 0);
      LMW_HIP(hipEventRecord(ev[q + 2], stream));
      if (mx) launch_anal_mx(ANAL_TEND, stage, dt, 0, B, nullptr);
      else launch_anal(ANAL_TEND, stage, dt, 0, B, nullptr);
      LMW_HIP(hipEventRecord(ev[q + 3], stream));
    }
  sync();
  for (size_t k = 0; k < ev.size(); k += 4)
    for (int j = 0; j < 3; j++) {
      float t = 0;
      LMW_HIP(hipEventElapsedTime(&t, ev[k + j], ev[k + j + 1]));
      ms[j] += t;
    }
  for (auto &e : ev) (void)hipEventDestroy(e);
}

void SphereCore::compute_grids(int b0, int nb) {
  launch_synth(x0.p, b0, nb);
  launch_to_grid(GRID_DIAG, grid.p, b0, nb);
}

void SphereCore::reduce(int b0, int nb) {
  hipLaunchKernelGGL(k_reduce_rows, dim3((J + 3) / 4, nb), dim3(256), 0, stream, geo(), grid.p, f.p, hs.p, ref_phi.p, rowred.p, b0, mflags.p);
  LMW_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_combine, dim3(nb), dim3(64), 0, stream, geo(), rowred.p, w.p, red.p, b0);
  LMW_LAUNCH_CHECK();
}

void SphereCore::download_grid(int b, int field, double *out) {
  const size_t plane = static_cast<size_t>(J) * K;
  grid.download(out, plane, (static_cast<size_t>(b) * 4 + field) * plane, stream);
  sync();
}

void SphereCore::download_red(int b, double *out16) {
  red.download(out16, 16, static_cast<size_t>(b) * 16, stream);
  sync();
}

double SphereCore::spectral_abs_sum(int b) {
  hipLaunchKernelGGL(k_abs_sum_member, dim3(1), dim3(256), 0, stream, geo(), b, x0.p, red.p + static_cast<size_t>(b) * 16 + 15);
  LMW_LAUNCH_CHECK();
  double s;
  red.download(&s, 1, static_cast<size_t>(b) * 16 + 15, stream);
  sync();
  return s;
}

void SphereCore::synthesize_raw(const double2 *z, const double2 *d, const double2 *phi, double *U, double *V, double *Z, double *PHI) {
  // member 0's Fourier and grid scratch; the input goes through spec_io laid out as one member's [3][S]
  std::vector<double2> zero(S, make_double2(0, 0));
  spec_io.upload(z ? z : zero.data(), S, 0, stream);
  spec_io.upload(d ? d : zero.data(), S, S, stream);
  spec_io.upload(phi ? phi : zero.data(), S, 2 * S, stream);
  launch_synth(spec_io.p, 0, 1, true);
  launch_to_grid(GRID_RAW, grid.p, 0, 1);
  const size_t plane = static_cast<size_t>(J) * K;
  if (U) grid.download(U, plane, 0, stream);
  if (V) grid.download(V, plane, plane, stream);
  if (Z) grid.download(Z, plane, 2 * plane, stream);
  if (PHI) grid.download(PHI, plane, 3 * plane, stream);
  sync();
}

void SphereCore::analyze_raw(int mode, const double *g0, const double *g1, const double *g2, double2 *out4) {
  const size_t plane = static_cast<size_t>(J) * K;
  if (g0) in3.upload(g0, plane, 0, stream);
  if (g1) in3.upload(g1, plane, plane, stream);
  if (g2) in3.upload(g2, plane, 2 * plane, stream);
  launch_prep(mode, 0, 1);
  launch_anal(ANAL_RAW, 0, 0.0, 0, 1, spec_io.p);
  spec_io.download(out4, 4 * static_cast<size_t>(S), 0, stream);
  sync();
}

void SphereCore::tendencies_raw(int b, double2 *dz, double2 *dd, double2 *df) {
  if (precision == LMW_FP64) {
    launch_synth(x0.p, b, 1);
    launch_tend_grid(b, 1, 0);
    launch_anal(ANAL_TENDOUT, 0, 0.0, b, 1, spec_io.p);
  } else {
    launch_synth_mx(x0.p, b, 1);
    launch_tend_grid(b, 1, 1);
    launch_anal_mx(ANAL_TENDOUT, 0, 0.0, b, 1, spec_io.p);
  }
  spec_io.download(dz, S, 0, stream);
  spec_io.download(dd, S, S, stream);
  spec_io.download(df, S, 2 * S, stream);
  sync();
}


// ------------------------------------------------------------------------------------------- free functions

// `count` independent in-place FFTs of length n in LDS (fft, sphere.cpp:21-45): one block per transform.
__global__ __launch_bounds__(256) void k_fft_batch(double2 *data, int n, int logn, const double2 *__restrict__ twg, int sign) {
  extern __shared__ double2 lds[];
  double2 *buf = lds, *tw = lds + n;
  double2 *row = data + static_cast<size_t>(blockIdx.x) * n;
  for (int i = threadIdx.x; i < n / 2; i += blockDim.x) tw[i] = twg[i];
  for (int i = threadIdx.x; i < n; i += blockDim.x) buf[brev(i, logn)] = row[i];
  __syncthreads();
  fft_lds(buf, 1, n, logn, tw, 0, sign);
  for (int i = threadIdx.x; i < n; i += blockDim.x) row[i] = buf[i];
}

// The same transforms through the grid kernel's four-step FFT (fft_lds.h), for the lengths the time step uses (the
// complex half-length L = K / 2 of its real rows, 8..256): natural-order input in LDS, fft2, output read back from
// its logical positions.  tw[q] = exp(-2 pi i q / (2 n)), q < 2 n, as the grid kernel's table.
template <int R1, int R2>
__global__ __launch_bounds__(256) void k_fft2_batch(double2 *data, const double2 *__restrict__ twg, int sign) {
  using F = Fft2<R1, R2>;
  constexpr int n = F::L;
  extern __shared__ double2 lds[];
  double2 *buf = lds, *tw = lds + F::P;
  double2 *row = data + static_cast<size_t>(blockIdx.x) * n;
  for (int i = threadIdx.x; i < 2 * n; i += blockDim.x) tw[i] = twg[i];
  for (int i = threadIdx.x; i < n; i += blockDim.x) buf[F::phys(i)] = row[i];
  __syncthreads();
  if (sign > 0) fft2<R1, R2, +1>(buf, 1, tw);
  else fft2<R1, R2, -1>(buf, 1, tw);
  for (int t = threadIdx.x; t < n; t += blockDim.x) row[t] = buf[F::phys(F::pos(t))];
}

void fft_batch(double2 *host, int n, int count, int sign) {
  if (n >= 8 && n <= 256) {  // the production four-step FFT
    std::vector<double2> htw(2 * n);
    for (int q = 0; q < 2 * n; q++) {
      const long double ang = 3.141592653589793238462643383279502884L * q / n;
      htw[q] = make_double2(static_cast<double>(cosl(ang)), static_cast<double>(-sinl(ang)));
    }
    DevArray<double2> d, t;
    d.alloc(static_cast<size_t>(n) * count);
    t.alloc(2 * n);
    LMW_HIP(hipMemcpy(d.p, host, d.bytes(), hipMemcpyHostToDevice));
    LMW_HIP(hipMemcpy(t.p, htw.data(), t.bytes(), hipMemcpyHostToDevice));
    const size_t pitch = static_cast<size_t>(n) + n / (n >= 64 ? 8 : n >= 16 ? 4 : 2);  // Fft2::P of the dispatched shape
    const size_t lds = (pitch + 2 * static_cast<size_t>(n)) * sizeof(double2);
    switch (n) {
      case 256: hipLaunchKernelGGL((k_fft2_batch<16, 16>), dim3(count), dim3(256), lds, 0, d.p, t.p, sign); break;
      case 128: hipLaunchKernelGGL((k_fft2_batch<16, 8>), dim3(count), dim3(256), lds, 0, d.p, t.p, sign); break;
      case 64: hipLaunchKernelGGL((k_fft2_batch<8, 8>), dim3(count), dim3(256), lds, 0, d.p, t.p, sign); break;
      case 32: hipLaunchKernelGGL((k_fft2_batch<8, 4>), dim3(count), dim3(256), lds, 0, d.p, t.p, sign); break;
      case 16: hipLaunchKernelGGL((k_fft2_batch<4, 4>), dim3(count), dim3(256), lds, 0, d.p, t.p, sign); break;
      default: hipLaunchKernelGGL((k_fft2_batch<4, 2>), dim3(count), dim3(256), lds, 0, d.p, t.p, sign); break;
    }
    LMW_LAUNCH_CHECK();
    LMW_HIP(hipMemcpy(host, d.p, d.bytes(), hipMemcpyDeviceToHost));
    return;
  }
  int logn = 0;
  while ((1 << logn) < n) logn++;
  std::vector<double2> htw(n / 2);
  for (int q = 0; q < n / 2; q++) {
    const long double ang = 2.0L * 3.141592653589793238462643383279502884L * q / n;
    htw[q] = make_double2(static_cast<double>(cosl(ang)), static_cast<double>(-sinl(ang)));
  }
  DevArray<double2> d, t;
  d.alloc(static_cast<size_t>(n) * count);
  t.alloc(n / 2);
  LMW_HIP(hipMemcpy(d.p, host, d.bytes(), hipMemcpyHostToDevice));
  LMW_HIP(hipMemcpy(t.p, htw.data(), t.bytes(), hipMemcpyHostToDevice));
  const size_t lds = (static_cast<size_t>(n) + n / 2) * sizeof(double2);
  hipLaunchKernelGGL(k_fft_batch, dim3(count), dim3(256), lds, 0, d.p, n, logn, t.p, sign);
  LMW_LAUNCH_CHECK();
  LMW_HIP(hipMemcpy(host, d.p, d.bytes(), hipMemcpyDeviceToHost));
}

void gauss_device(int n, double *x, double *w) {
  DevArray<double> dx, dw;
  dx.alloc(n);
  dw.alloc(n);
  hipLaunchKernelGGL(k_gauss, dim3((n + 255) / 256), dim3(256), 0, 0, n, dx.p, dw.p);
  LMW_LAUNCH_CHECK();
  LMW_HIP(hipMemcpy(x, dx.p, dx.bytes(), hipMemcpyDeviceToHost));
  LMW_HIP(hipMemcpy(w, dw.p, dw.bytes(), hipMemcpyDeviceToHost));
}

}  // namespace lmw
