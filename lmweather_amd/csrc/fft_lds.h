// Four-step FFT of length L = R1 * R2 in LDS with the small DFTs in registers (radix 2..16), for the per-latitude
// Fourier transforms of the sphere (reference fft, cpp/sphere.cpp:21-45, applied per row in :123-136).
//
//   pass 1, per n2 < R2:  y[n2][k1] = W_L^(n2 k1) sum_n1 x[R2 n1 + n2] W_R1^(n1 k1)      (in place)
//   pass 2, per k1 < R1:  X[k1 + R1 k2] = sum_n2 y[n2][k1] W_R2^(n2 k2)                  (in place)
//
// Input is in natural order; output index t lands at logical position pos(t) = R2 (t mod R1) + t div R1.  Logical
// position p lives at LDS word phys(p) = p + p / R2 (one pad per R2 entries), so that both passes -- pass 1 reading
// stride R2 across a thread's registers, pass 2 reading rows of R2 -- hit distinct LDS banks across a wavefront.
// Two synchronizations per transform instead of log2(L) for radix 2.  tw[q] = exp(-2 pi i q / (2L)), q < 2L.
#pragma once
#include "common.h"

namespace lmw {

template <int R1, int R2> struct Fft2 {
  static constexpr int L = R1 * R2;
  static constexpr int P = L + L / R2;  // padded pitch of one array
  static constexpr int LOG2R2 = R2 == 2 ? 1 : R2 == 4 ? 2 : R2 == 8 ? 3 : 4;
  __device__ static inline int phys(int p) { return p + (p >> LOG2R2); }
  __device__ static inline int pos(int t) { return R2 * (t % R1) + t / R1; }
};

namespace fftd {
constexpr int brev_c(int i, int R) {
  int r = 0;
  for (int b = R >> 1, s = 1; b; b >>= 1, s <<= 1)
    if (i & s) r |= b;
  return r;
}
// W_16^k = exp(-2 pi i k / 16) for k < 8 (cos and sin of k pi / 8, correctly rounded)
__device__ inline double2 w16(int k) {
  constexpr double c1 = 0.9238795325112867, s1 = 0.3826834323650898, h = 0.7071067811865476;
  switch (k & 7) {
    case 0: return make_double2(1.0, 0.0);
    case 1: return make_double2(c1, -s1);
    case 2: return make_double2(h, -h);
    case 3: return make_double2(s1, -c1);
    case 4: return make_double2(0.0, -1.0);
    case 5: return make_double2(-s1, -c1);
    case 6: return make_double2(-h, -h);
    default: return make_double2(-c1, -s1);
  }
}
}  // namespace fftd

// v * W_16^K (SIGN < 0) or v * conj(W_16^K) (SIGN > 0), with the multiplications by 1, -i, +i and (1 -+ i)/sqrt 2
// done as swaps, negations and one scaled sum/difference instead of a full complex product.  C = double2 or float2.
template <int K, int SIGN, class C> __device__ inline C mul_w16(C v) {
  using T = typename real_of<C>::type;
  constexpr T h = static_cast<T>(0.7071067811865476);
  constexpr int k = K & 7;
  if constexpr (k == 0) {
    return v;
  } else if constexpr (k == 4) {
    return SIGN < 0 ? C{v.y, -v.x} : C{-v.y, v.x};  // * -i, * +i
  } else if constexpr (k == 2) {
    return SIGN < 0 ? C{h * (v.x + v.y), h * (v.y - v.x)} : C{h * (v.x - v.y), h * (v.x + v.y)};
  } else if constexpr (k == 6) {
    return SIGN < 0 ? C{h * (v.y - v.x), -h * (v.x + v.y)} : C{-h * (v.x + v.y), h * (v.x - v.y)};
  } else {
    const double2 wd = fftd::w16(k);
    C w = mkc<C>(wd.x, wd.y);
    if (SIGN > 0) w.y = -w.y;
    return cmul(v, w);
  }
}

template <int R, int SIGN, int LEN, int I, int KK, class C> __device__ inline void butterflies(C (&v)[R]) {
  if constexpr (KK < LEN / 2) {
    const C t = mul_w16<KK * (16 / LEN), SIGN>(v[I + KK + LEN / 2]);
    v[I + KK + LEN / 2] = csub(v[I + KK], t);
    v[I + KK] = cadd(v[I + KK], t);
    butterflies<R, SIGN, LEN, I, KK + 1>(v);
  }
}
template <int R, int SIGN, int LEN, int I, class C> __device__ inline void stage(C (&v)[R]) {
  if constexpr (I < R) {
    butterflies<R, SIGN, LEN, I, 0>(v);
    stage<R, SIGN, LEN, I + LEN>(v);
  }
}
template <int R, int SIGN, int LEN, class C> __device__ inline void stages(C (&v)[R]) {
  if constexpr (LEN <= R) {
    stage<R, SIGN, LEN, 0>(v);
    stages<R, SIGN, 2 * LEN>(v);
  }
}

// In-register DFT of length R (<= 16): v_k <- sum_n v_n exp(SIGN 2 pi i n k / R), radix-2 decimation in time.
template <int R, int SIGN, class C> __device__ inline void dft_reg(C (&v)[R]) {
#pragma unroll
  for (int i = 0; i < R; i++) {
    const int j = fftd::brev_c(i, R);
    if (i < j) {
      const C t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
  stages<R, SIGN, 2>(v);
}

// nf arrays of pitch Fft2::P at buf, transformed in place; all threads of the block take part.
template <int R1, int R2, int SIGN, class C> __device__ inline void fft2(C *buf, int nf, const C *tw) {
  using F = Fft2<R1, R2>;
  for (int it = threadIdx.x; it < nf * R2; it += blockDim.x) {
    const int f = it / R2, n2 = it - f * R2;
    C *a = buf + f * F::P;
    C v[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++) v[n1] = a[F::phys(R2 * n1 + n2)];
    dft_reg<R1, SIGN>(v);
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) {
      C w = tw[2 * n2 * k1];
      if (SIGN > 0) w.y = -w.y;
      a[F::phys(R2 * k1 + n2)] = cmul(v[k1], w);
    }
  }
  __syncthreads();
  for (int it = threadIdx.x; it < nf * R1; it += blockDim.x) {
    const int f = it / R1, k1 = it - f * R1;
    C *a = buf + f * F::P;
    C v[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; n2++) v[n2] = a[F::phys(R2 * k1 + n2)];
    dft_reg<R2, SIGN>(v);
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) a[F::phys(R2 * k1 + k2)] = v[k2];
  }
  __syncthreads();
}

}  // namespace lmw
