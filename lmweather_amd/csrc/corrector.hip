// Learned coarse-step corrector: a small spectral-space network, trained and run on the GPU with hand-written
// forward, backward and Adam kernels (BF16 weights for the arithmetic, FP32 master weights and FP32 accumulation).
//
// The coarse step integrates dt as `coarse_substeps` RK4 steps of the spectral model (a single RK4 step of
// 2400 s at T42 is unstable: the fastest gravity mode has |omega dt| near 4.4, where |R(i y)| of RK4 is about 12),
// and the network adds a correction toward the high-resolution model truncated to the same triangle.
//
// Rows are (sample, packed coefficient (m, n)), C = (N+1)(N+2)/2 rows per sample, packed index
// m (N+1) - m (m-1)/2 + n - m (the layout of the model state).  Row features (FIN = 8): real and imaginary parts of
// zeta, delta, phi at (m, n), each divided by a per-degree input scale, then n/N and m/N.  Four dense layers of
// width W act on every row with shared weights; the two inner layers also see the mean over m of the previous
// layer's activations at the row's degree n (per-degree mixing, the only place rows of a sample meet):
//
//   H1 = tanh(X W1^T + b1)
//   H2 = tanh(H1 W2^T + b2 + mean_m(H1)[n] V2^T)
//   H3 = tanh(H2 W3^T + b3 + mean_m(H2)[n] V3^T)
//   Y  = H3 W4^T + b4                                     (FOUT = 6 outputs, in per-degree output-scale units)
//
// Every product uses the BF16-rounded copy of the weights and FP32 accumulation; Adam updates the FP32 master
// copy.  Loss: mean squared error over the trained entries, which exclude the (0, 0) row (the global means of
// zeta, delta, phi: a correction there would break conservation) and the imaginary parts at m = 0 (zero for a
// real field).  The correction is zero on those entries.
#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <vector>

#include "sphere_core.h"

namespace lmw {

SphereCore &core_of(lmw_sphere *s);
SphereCore &single_group(lmw_ensemble *e);

namespace {
constexpr int FIN = 8, FOUT = 6, WMAX = 64;
constexpr int WROWS = 256;  // rows per block of the weight-gradient partial sums

__device__ inline float bf16r(float x) { return static_cast<float>(static_cast<__bf16>(x)); }
__host__ __device__ inline int tri_off(int m, int N1) { return m * N1 - m * (m - 1) / 2; }
// (m, n) of packed index c in a triangle of N1 = N + 1 orders
__device__ inline void tri_mn(int c, int N1, int &m, int &n) {
  int mm = 0;
  while (mm + 1 < N1 && tri_off(mm + 1, N1) <= c) mm++;
  m = mm;
  n = c - tri_off(mm, N1) + mm;
}

// Y[r][o] = act(sum_i X[r][i] W[o][i] + b[o] + mix[(r / C) * N1 + deg[r % C]][o]).
// One block per 64 rows; thread t computes row t % 64 for a quarter of the outputs.  I, O <= 64.
__global__ __launch_bounds__(256) void k_dense_fwd(const float *__restrict__ X, int R, int I, const float *__restrict__ Wb,
                                                   const float *__restrict__ b, int O, const float *__restrict__ mix, const int *__restrict__ deg,
                                                   int C, int N1, int act, float *__restrict__ Y) {
  __shared__ float xs[64][WMAX + 1];
  __shared__ float ws[WMAX][WMAX + 1];
  const int r0 = blockIdx.x * 64, t = threadIdx.x;
  for (int e = t; e < 64 * I; e += 256) {
    const int rr = e / I, i = e % I;
    xs[rr][i] = (r0 + rr < R) ? X[static_cast<size_t>(r0 + rr) * I + i] : 0.f;
  }
  for (int e = t; e < O * I; e += 256) ws[e / I][e % I] = Wb[e];
  __syncthreads();
  const int rr = t & 63, og = t >> 6, r = r0 + rr;
  if (r >= R) return;
  const int oq = (O + 3) / 4;
  for (int o = og * oq; o < min(O, (og + 1) * oq); o++) {
    float acc = b ? b[o] : 0.f;
    for (int i = 0; i < I; i++) acc = fmaf(xs[rr][i], ws[o][i], acc);
    if (mix) acc += mix[(static_cast<size_t>(r / C) * N1 + deg[r % C]) * O + o];
    Y[static_cast<size_t>(r) * O + o] = act ? tanhf(acc) : acc;
  }
}

// ---------------------------------------------------------------- dense layers on the BF16 MFMA (16x16x32)
// The weights used for the arithmetic are BF16 values already (wb); each FP32 activation or gradient x is split into
// three BF16 parts x1 + x2 + x3 = x exactly, so the three products w x1 + w x2 + w x3 are exact and the FP32
// accumulation gives the FP32 sum of the exact products (in the MFMA's order).  Lane l supplies A[l % 16][8 (l / 16)
// .. + 7] and B[8 (l / 16) .. + 7][l % 16]; it receives D[4 (l / 16) + i][l % 16] in register i.
typedef float mx_v4f __attribute__((ext_vector_type(4)));
typedef __bf16 mx_v8bf __attribute__((ext_vector_type(8)));

__device__ inline void split3(const float (&x)[8], mx_v8bf &a1, mx_v8bf &a2, mx_v8bf &a3) {
#pragma unroll
  for (int t = 0; t < 8; t++) {
    const __bf16 h = static_cast<__bf16>(x[t]);
    const float r1 = x[t] - static_cast<float>(h);
    const __bf16 m = static_cast<__bf16>(r1);
    a1[t] = h;
    a2[t] = m;
    a3[t] = static_cast<__bf16>(r1 - static_cast<float>(m));
  }
}
__device__ inline mx_v4f mma3x(const mx_v8bf &a1, const mx_v8bf &a2, const mx_v8bf &a3, const mx_v8bf &w, mx_v4f c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, w, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, w, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, w, c, 0, 0, 0);
}

// Y[r][o] = act(sum_i X[r][i] W[o][i] + b[o] + mix[(r / C) * N1 + deg[r % C]][o]) for I a multiple of 32 and
// O <= 16 NCT.  Block: 4 waves, one 16-row tile each; a wave computes all NCT column tiles of its rows.
template <int NCT>
__global__ __launch_bounds__(256) void k_dense_fwd_mx(const float *__restrict__ X, int R, int I, const float *__restrict__ Wb,
                                                      const float *__restrict__ b, int O, const float *__restrict__ mix,
                                                      const int *__restrict__ deg, int C, int N1, int act, float *__restrict__ Y) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r0 = (blockIdx.x * 4 + wave) * 16;
  if (r0 >= R) return;
  const int c16 = lane & 15, kq = lane >> 4;
  mx_v4f acc[NCT];
#pragma unroll
  for (int t = 0; t < NCT; t++) acc[t] = mx_v4f{0, 0, 0, 0};
  const int ra = r0 + c16;
  for (int k0 = 0; k0 < I; k0 += 32) {
    float x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (ra < R) {
      const float4 *p = reinterpret_cast<const float4 *>(X + static_cast<size_t>(ra) * I + k0 + 8 * kq);
      const float4 u = p[0], v = p[1];
      x[0] = u.x, x[1] = u.y, x[2] = u.z, x[3] = u.w, x[4] = v.x, x[5] = v.y, x[6] = v.z, x[7] = v.w;
    }
    mx_v8bf x1, x2, x3;
    split3(x, x1, x2, x3);
#pragma unroll
    for (int t = 0; t < NCT; t++) {
      const int o = 16 * t + c16;
      mx_v8bf w;
      if (o < O) {
        const float4 *q = reinterpret_cast<const float4 *>(Wb + static_cast<size_t>(o) * I + k0 + 8 * kq);
        const float4 u = q[0], v = q[1];
        w = mx_v8bf{static_cast<__bf16>(u.x), static_cast<__bf16>(u.y), static_cast<__bf16>(u.z), static_cast<__bf16>(u.w),
                    static_cast<__bf16>(v.x), static_cast<__bf16>(v.y), static_cast<__bf16>(v.z), static_cast<__bf16>(v.w)};
      } else {
        w = mx_v8bf{0, 0, 0, 0, 0, 0, 0, 0};
      }
      acc[t] = mma3x(x1, x2, x3, w, acc[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < NCT; t++) {
    const int o = 16 * t + c16;
    if (o >= O) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int r = r0 + 4 * kq + i;
      if (r >= R) continue;
      float v = acc[t][i] + (b ? b[o] : 0.f);
      if (mix) v += mix[(static_cast<size_t>(r / C) * N1 + deg[r % C]) * O + o];
      Y[static_cast<size_t>(r) * O + o] = act ? tanhf(v) : v;
    }
  }
}

// dX[r][i] = sum_o dZ[r][o] W[o][i] for I a multiple of 16 (I / 16 <= NCT column tiles) and O <= 64: W transposed
// into LDS as BF16 [i][o], o padded to a multiple of 32 with zeros.
template <int NCT>
__global__ __launch_bounds__(256) void k_dense_bwd_data_mx(const float *__restrict__ dZ, int R, int O, const float *__restrict__ Wb, int I,
                                                           float *__restrict__ dX) {
  __shared__ __bf16 wt[16 * NCT][WMAX + 8];
  const int OP = (O + 31) / 32 * 32;
  for (int e = threadIdx.x; e < 16 * NCT * OP; e += blockDim.x) {
    const int i = e / OP, o = e % OP;
    wt[i][o] = static_cast<__bf16>(i < I && o < O ? Wb[static_cast<size_t>(o) * I + i] : 0.f);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r0 = (blockIdx.x * 4 + wave) * 16;
  if (r0 >= R) return;
  const int c16 = lane & 15, kq = lane >> 4, ra = r0 + c16;
  mx_v4f acc[NCT];
#pragma unroll
  for (int t = 0; t < NCT; t++) acc[t] = mx_v4f{0, 0, 0, 0};
  for (int k0 = 0; k0 < OP; k0 += 32) {
    float x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int o = k0 + 8 * kq + u;
      x[u] = ra < R && o < O ? dZ[static_cast<size_t>(ra) * O + o] : 0.f;
    }
    mx_v8bf x1, x2, x3;
    split3(x, x1, x2, x3);
#pragma unroll
    for (int t = 0; t < NCT; t++) {
      const mx_v8bf w = *reinterpret_cast<const mx_v8bf *>(&wt[16 * t + c16][k0 + 8 * kq]);
      acc[t] = mma3x(x1, x2, x3, w, acc[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < NCT; t++) {
    const int i = 16 * t + c16;
    if (i >= I) continue;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int r = r0 + 4 * kq + q;
      if (r < R) dX[static_cast<size_t>(r) * I + i] = acc[t][q];
    }
  }
}

// The weight-gradient partial sums of k_dense_bwd_weight on the BF16 MFMA: part[blk][o I + i] = sum over the WROWS
// rows of block blk of dZ[r][o] X[r][i], part[blk][O I + o] = sum dZ[r][o].  Both operands are FP32 and split into
// three BF16 parts; the six products x1 y1 + x1 y2 + x2 y1 + x1 y3 + x2 y2 + x3 y1 keep FP32 accuracy (the dropped
// terms are below 2^-24 relative).  Rows are staged 32 at a time in LDS, transposed to [o][r] and [i][r], so that a
// lane's 8 consecutive k (rows) are contiguous.  Block: 4 waves over the (O / 16) x (I / 16) output tiles.
__global__ __launch_bounds__(256) void k_dense_bwd_weight_mx(const float *__restrict__ dZ, const float *__restrict__ X, int R, int O, int I,
                                                             float *__restrict__ part) {
  constexpr int RC = 32, PT = RC + 8;  // rows per chunk, padded pitch (bf16)
  __shared__ __bf16 zt[3][WMAX][PT], xt[3][WMAX][PT];
  __shared__ float bsum[4][WMAX];
  const int blk = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rb = blk * WROWS, re = min(R, rb + WROWS);
  const int OT = (O + 15) / 16, IT = (I + 15) / 16, ntile = OT * IT;
  mx_v4f acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = mx_v4f{0, 0, 0, 0};
  float bs = 0;  // thread t < O: the bias sum of output t over this wave's rows of the chunk (4 waves, combined below)
  for (int c0 = rb; c0 < re; c0 += RC) {
    __syncthreads();
    for (int e = threadIdx.x; e < RC * WMAX; e += blockDim.x) {
      const int rr = e / WMAX, c = e % WMAX, r = c0 + rr;
      const float z = r < re && c < O ? dZ[static_cast<size_t>(r) * O + c] : 0.f;
      const float x = r < re && c < I ? X[static_cast<size_t>(r) * I + c] : 0.f;
      const __bf16 z1 = static_cast<__bf16>(z), x1 = static_cast<__bf16>(x);
      const float zr = z - static_cast<float>(z1), xr = x - static_cast<float>(x1);
      const __bf16 z2 = static_cast<__bf16>(zr), x2 = static_cast<__bf16>(xr);
      zt[0][c][rr] = z1;
      zt[1][c][rr] = z2;
      zt[2][c][rr] = static_cast<__bf16>(zr - static_cast<float>(z2));
      xt[0][c][rr] = x1;
      xt[1][c][rr] = x2;
      xt[2][c][rr] = static_cast<__bf16>(xr - static_cast<float>(x2));
    }
    __syncthreads();
    if (threadIdx.x < O)
      for (int rr = 0; rr < RC && c0 + rr < re; rr++) bs += dZ[static_cast<size_t>(c0 + rr) * O + threadIdx.x];
    const int kk = 8 * (lane >> 4), c16 = lane & 15;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int t = wave + 4 * q;
      if (t >= ntile) break;
      const int o = 16 * (t / IT) + c16, i = 16 * (t % IT) + c16;
      const mx_v8bf a1 = *reinterpret_cast<const mx_v8bf *>(&zt[0][o][kk]), a2 = *reinterpret_cast<const mx_v8bf *>(&zt[1][o][kk]),
                    a3 = *reinterpret_cast<const mx_v8bf *>(&zt[2][o][kk]);
      const mx_v8bf b1 = *reinterpret_cast<const mx_v8bf *>(&xt[0][i][kk]), b2 = *reinterpret_cast<const mx_v8bf *>(&xt[1][i][kk]),
                    b3 = *reinterpret_cast<const mx_v8bf *>(&xt[2][i][kk]);
      mx_v4f c = acc[q];
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, b1, c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b2, c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b3, c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b1, c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b2, c, 0, 0, 0);
      acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, c, 0, 0, 0);
    }
  }
  const int nw = O * I + O;
  float *out = part + static_cast<size_t>(blk) * nw;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int t = wave + 4 * q;
    if (t >= ntile) break;
    const int i = 16 * (t % IT) + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int o = 16 * (t / IT) + 4 * (lane >> 4) + r;
      if (o < O && i < I) out[static_cast<size_t>(o) * I + i] = acc[q][r];
    }
  }
  if (threadIdx.x < O) out[O * I + threadIdx.x] = bs;
  (void)bsum;
}

// dX[r][i] = sum_o dZ[r][o] W[o][i]
__global__ __launch_bounds__(256) void k_dense_bwd_data(const float *__restrict__ dZ, int R, int O, const float *__restrict__ Wb, int I,
                                                        float *__restrict__ dX) {
  __shared__ float ds[64][WMAX + 1];
  __shared__ float ws[WMAX][WMAX + 1];
  const int r0 = blockIdx.x * 64, t = threadIdx.x;
  for (int e = t; e < 64 * O; e += 256) {
    const int rr = e / O, o = e % O;
    ds[rr][o] = (r0 + rr < R) ? dZ[static_cast<size_t>(r0 + rr) * O + o] : 0.f;
  }
  for (int e = t; e < O * I; e += 256) ws[e / I][e % I] = Wb[e];
  __syncthreads();
  const int rr = t & 63, ig = t >> 6, r = r0 + rr;
  if (r >= R) return;
  const int iq = (I + 3) / 4;
  for (int i = ig * iq; i < min(I, (ig + 1) * iq); i++) {
    float acc = 0;
    for (int o = 0; o < O; o++) acc = fmaf(ds[rr][o], ws[o][i], acc);
    dX[static_cast<size_t>(r) * I + i] = acc;
  }
}

// Partial sums over the WROWS rows of block blk: part[blk][o * I + i] = sum dZ[r][o] X[r][i], part[blk][O*I + o] =
// sum dZ[r][o].  Thread t owns entries t, t + 256, ... (at most 17: O*I + O <= 4160).
__global__ __launch_bounds__(256) void k_dense_bwd_weight(const float *__restrict__ dZ, const float *__restrict__ X, int R, int O, int I,
                                                          float *__restrict__ part) {
  __shared__ float ds[32][WMAX + 1];
  __shared__ float xs[32][WMAX + 1];
  const int blk = blockIdx.x, t = threadIdx.x;
  const int rb = blk * WROWS, re = min(R, rb + WROWS);
  const int nw = O * I + O;
  float acc[17];
#pragma unroll
  for (int q = 0; q < 17; q++) acc[q] = 0;
  for (int c0 = rb; c0 < re; c0 += 32) {
    __syncthreads();
    for (int e = t; e < 32 * O; e += 256) {
      const int rr = e / O, o = e % O;
      ds[rr][o] = (c0 + rr < re) ? dZ[static_cast<size_t>(c0 + rr) * O + o] : 0.f;
    }
    for (int e = t; e < 32 * I; e += 256) {
      const int rr = e / I, i = e % I;
      xs[rr][i] = (c0 + rr < re) ? X[static_cast<size_t>(c0 + rr) * I + i] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 17; q++) {
      const int w = t + 256 * q;
      if (w < nw) {
        float s = 0;
        if (w < O * I) {
          const int o = w / I, i = w % I;
          for (int rr = 0; rr < 32; rr++) s = fmaf(ds[rr][o], xs[rr][i], s);
        } else {
          const int o = w - O * I;
          for (int rr = 0; rr < 32; rr++) s += ds[rr][o];
        }
        acc[q] += s;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 17; q++) {
    const int w = t + 256 * q;
    if (w < nw) part[static_cast<size_t>(blk) * nw + w] = acc[q];
  }
}

// gradient = sum of the partials in block order (deterministic); weights to gw, biases (if gb) to gb
__global__ void k_sum_parts(const float *__restrict__ part, int nblk, int O, int I, float *__restrict__ gw, float *__restrict__ gb) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x, nw = O * I + O;
  if (w >= nw) return;
  float s = 0;
  for (int k = 0; k < nblk; k++) s += part[static_cast<size_t>(k) * nw + w];
  if (w < O * I) gw[w] = s;
  else if (gb) gb[w - O * I] = s;
}

// per-degree mean over m of H: P[s][n][f] = mean over m <= n of H[s][(m, n)][f]
__global__ void k_pool(const float *__restrict__ H, int S, int C, int F, int N1, float *__restrict__ P) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= S * N1 * F) return;
  const int f = idx % F, n = (idx / F) % N1, s = idx / (F * N1);
  float acc = 0;
  for (int m = 0; m <= n; m++) acc += H[(static_cast<size_t>(s) * C + tri_off(m, N1) + n - m) * F + f];
  P[idx] = acc / (n + 1);
}

// per-degree sums of dZ over m (for the mixing weights): Q[s][n][f]
__global__ void k_degsum(const float *__restrict__ dZ, int S, int C, int F, int N1, float *__restrict__ Q) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= S * N1 * F) return;
  const int f = idx % F, n = (idx / F) % N1, s = idx / (F * N1);
  float acc = 0;
  for (int m = 0; m <= n; m++) acc += dZ[(static_cast<size_t>(s) * C + tri_off(m, N1) + n - m) * F + f];
  Q[idx] = acc;
}

// dH[r][f] += G[s][deg][f] / (deg + 1): the gradient through the mean over m
__global__ void k_unpool_add(const float *__restrict__ Gm, int S, int C, int F, int N1, const int *__restrict__ deg, float *__restrict__ dH) {
  const size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<size_t>(S) * C * F) return;
  const int f = static_cast<int>(idx % F), c = static_cast<int>((idx / F) % C), s = static_cast<int>(idx / (static_cast<size_t>(F) * C));
  const int n = deg[c];
  dH[idx] += Gm[(static_cast<size_t>(s) * N1 + n) * F + f] / (n + 1);
}

// d *= 1 - H^2 (tanh')
__global__ void k_tanh_bwd(const float *__restrict__ H, float *__restrict__ d, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) d[i] *= 1.f - H[i] * H[i];
}

// normalized features (F = FIN: with n/N, m/N) or targets (F = FOUT) from raw spectra [s][3][C][2]
__global__ void k_features(const float *__restrict__ raw, int S, int C, int N, const int *__restrict__ deg, const int *__restrict__ ordm,
                           const float *__restrict__ scale, int F, float *__restrict__ X) {
  const size_t r = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= static_cast<size_t>(S) * C) return;
  const int c = static_cast<int>(r % C), s = static_cast<int>(r / C), n = deg[c], m = ordm[c];
  for (int f = 0; f < 3; f++) {
    const float sc = scale[f * (N + 1) + n];
    const size_t src = ((static_cast<size_t>(s) * 3 + f) * C + c) * 2;
    X[r * F + 2 * f] = raw[src] / sc;
    X[r * F + 2 * f + 1] = raw[src + 1] / sc;
  }
  if (F == FIN) {
    X[r * F + 6] = static_cast<float>(n) / N;
    X[r * F + 7] = static_cast<float>(m) / N;
  }
}

__device__ inline bool masked_entry(int c, int m, int k) { return c == 0 || (m == 0 && (k & 1)); }

// squared errors per row; dY = 2 (Y - T) / count on trained entries (0 on masked ones)
__global__ void k_loss(const float *__restrict__ Y, const float *__restrict__ T, int S, int C, const int *__restrict__ ordm, float inv_count,
                       float *__restrict__ dY, float *__restrict__ lossrow) {
  const size_t r = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= static_cast<size_t>(S) * C) return;
  const int c = static_cast<int>(r % C), m = ordm[c];
  float l = 0;
  for (int k = 0; k < FOUT; k++) {
    const float d = masked_entry(c, m, k) ? 0.f : Y[r * FOUT + k] - T[r * FOUT + k];
    l = fmaf(d, d, l);
    if (dY) dY[r * FOUT + k] = 2.f * d * inv_count;
  }
  lossrow[r] = l;
}

// out[slot] = sum of v[0..n) (one block, FP64 accumulation)
__global__ __launch_bounds__(256) void k_sum_rows(const float *__restrict__ v, size_t n, double *__restrict__ out, int slot) {
  __shared__ double red[256];
  double s = 0;
  for (size_t i = threadIdx.x; i < n; i += 256) s += v[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[slot] = red[0];
}

// sum of squares of the trained outputs of each sample (network units): one block per sample; with per > 0 the rms
// sqrt(sum / per) instead
__global__ __launch_bounds__(256) void k_member_sq(const float *__restrict__ Y, int C, const int *__restrict__ ordm, double per,
                                                   double *__restrict__ res) {
  __shared__ double red[256];
  const int s = blockIdx.x;
  double acc = 0;
  for (int e = threadIdx.x; e < C * FOUT; e += 256) {
    const int c = e / FOUT, k = e % FOUT;
    if (!masked_entry(c, ordm[c], k)) {
      const double y = Y[static_cast<size_t>(s) * C * FOUT + e];
      acc += y * y;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) res[s] = per > 0 ? sqrt(red[0] / per) : red[0];
}

// correction in raw units: out[s][f][c][h] = Y[s][c][2f+h] * out_scale[f][n], zero on masked entries
__global__ void k_unnormalize(const float *__restrict__ Y, int S, int C, int N, const int *__restrict__ deg, const int *__restrict__ ordm,
                              const float *__restrict__ scale, float *__restrict__ out) {
  const size_t r = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= static_cast<size_t>(S) * C) return;
  const int c = static_cast<int>(r % C), s = static_cast<int>(r / C), n = deg[c], m = ordm[c];
  for (int f = 0; f < 3; f++)
    for (int h = 0; h < 2; h++)
      out[((static_cast<size_t>(s) * 3 + f) * C + c) * 2 + h] = masked_entry(c, m, 2 * f + h) ? 0.f : Y[r * FOUT + 2 * f + h] * scale[f * (N + 1) + n];
}

// Adam on the FP32 master weights, then their BF16-rounded compute copy
__global__ void k_adam(float *__restrict__ w, float *__restrict__ wb, const float *__restrict__ g, float *__restrict__ m1, float *__restrict__ m2,
                       size_t n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  const float nw = w[i] - lr * (a / bc1) / (sqrtf(v / bc2) + eps);
  w[i] = nw;
  wb[i] = bf16r(nw);
}

__global__ void k_round_bf16(const float *__restrict__ w, float *__restrict__ wb, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = bf16r(w[i]);
}

// dst[q] = src[idx[q]] for `count` samples of `per` floats each
__global__ void k_gather(const float *__restrict__ src, const int *__restrict__ idx, int count, size_t per, float *__restrict__ dst) {
  const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= per * count) return;
  const size_t q = e / per, k = e % per;
  dst[e] = src[static_cast<size_t>(idx[q]) * per + k];
}

// the state of every member (layout G, triangle of N + 1 orders) truncated to T, as float [b][3][C_T][2]
__global__ void k_truncate_out(Geo G, const double2 *__restrict__ x, int T, float *__restrict__ out) {
  const int T1 = T + 1, CT = T1 * (T1 + 1) / 2, N1f = G.N + 1;
  const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<size_t>(G.B) * 3 * CT) return;
  const int c = static_cast<int>(e % CT);
  const size_t bf = e / CT;
  int m, n;
  tri_mn(c, T1, m, n);
  const double2 v = x[G.at(static_cast<int>(bf / 3), static_cast<int>(bf % 3), tri_off(m, N1f) + n - m)];
  out[2 * e] = static_cast<float>(v.x);
  out[2 * e + 1] = static_cast<float>(v.y);
}

// the inverse: every coefficient of every member set from float [b][3][C_T][2] (zero above T), in x0 and xs
__global__ void k_truncate_in(Geo G, const float *__restrict__ in, int T, double2 *__restrict__ x0, double2 *__restrict__ xs) {
  const int T1 = T + 1, CT = T1 * (T1 + 1) / 2, N1f = G.N + 1, S = G.S;
  const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<size_t>(G.B) * 3 * S) return;
  const int c = static_cast<int>(e % S);
  const size_t bf = e / S;
  int m, n;
  tri_mn(c, N1f, m, n);
  double2 v = make_double2(0, 0);
  if (n <= T) {
    const size_t i = bf * CT + tri_off(m, T1) + n - m;
    v = make_double2(in[2 * i], in[2 * i + 1]);
  }
  const size_t o = G.at(static_cast<int>(bf / 3), static_cast<int>(bf % 3), c);
  x0[o] = v;
  xs[o] = v;
}

// x0 += correction (float [b][3][S][2]) and xs = x0, for the members with use[b] != 0
__global__ void k_add_correction(Geo G, double2 *__restrict__ x0, double2 *__restrict__ xs, const float *__restrict__ corr, const int *__restrict__ use) {
  const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t S = G.S;
  if (e >= static_cast<size_t>(G.B) * 3 * S) return;
  const int b = static_cast<int>(e / (3 * S)), f = static_cast<int>((e / S) % 3);
  if (!use[b]) return;
  const size_t o = G.at(b, f, e % S);
  const double2 v = make_double2(x0[o].x + corr[2 * e], x0[o].y + corr[2 * e + 1]);
  x0[o] = v;
  xs[o] = v;
}

// x0 = xs = kept for the members with use[b] != 0 (every array in layout G)
__global__ void k_restore_members(Geo G, const double2 *__restrict__ kept, const int *__restrict__ use, double2 *__restrict__ x0,
                                  double2 *__restrict__ xs) {
  const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t S = G.S;
  if (e >= static_cast<size_t>(G.B) * 3 * S) return;
  const int b = static_cast<int>(e / (3 * S));
  if (!use[b]) return;
  const size_t o = G.at(b, static_cast<int>((e / S) % 3), e % S);
  x0[o] = kept[o];
  xs[o] = kept[o];
}

unsigned blocks(size_t n, unsigned t = 256) { return static_cast<unsigned>((n + t - 1) / t); }
}  // namespace

// ------------------------------------------------------------------------------------------------ host side

struct Corrector {
  int N, N1, C, W;
  // flat parameter vector: W1 [W][8], b1 [W], W2 [W][W], b2 [W], V2 [W][W], W3 [W][W], b3 [W], V3 [W][W], W4 [6][W], b4 [6]
  size_t oW1, ob1, oW2, ob2, oV2, oW3, ob3, oV3, oW4, ob4, np;
  DevArray<float> w, wb, g, m1, m2;
  DevArray<int> deg, ordm;
  DevArray<float> in_scale, out_scale;
  int64_t adam_t = 0;
  hipStream_t stream = nullptr;
  // workspace for up to `cap` samples
  int cap = 0;
  DevArray<float> X, H1, H2, H3, Y, T, dY, dA, dB, P1, P2, PM, Q, Gm, part, lossrow, raw, raw2;
  DevArray<double> sums;
  // resident training set: inputs (states after the coarse step) and target corrections, raw [n][3][C][2]
  DevArray<float> xs, ts;
  DevArray<int> perm;
  int nsamples = 0;

  Corrector(int truncation, int width, uint64_t seed) : N(truncation), N1(truncation + 1), W(width) {
    if (truncation < 4 || truncation > 170) invalid("truncation must be between 4 and 170");
    if (width < 8 || width > WMAX) invalid("corrector width must be between 8 and 64");
    C = N1 * (N1 + 1) / 2;
    size_t o = 0;
    auto take = [&](size_t n) {
      const size_t r = o;
      o += n;
      return r;
    };
    const size_t WW = static_cast<size_t>(W) * W;
    oW1 = take(static_cast<size_t>(W) * FIN);
    ob1 = take(W);
    oW2 = take(WW);
    ob2 = take(W);
    oV2 = take(WW);
    oW3 = take(WW);
    ob3 = take(W);
    oV3 = take(WW);
    oW4 = take(static_cast<size_t>(FOUT) * W);
    ob4 = take(FOUT);
    np = o;
    LMW_HIP(hipStreamCreate(&stream));
    w.alloc(np);
    wb.alloc(np);
    g.alloc(np);
    m1.alloc(np);
    m2.alloc(np);
    // Glorot-uniform initialization from splitmix64; the mixing and output layers start small so the initial
    // correction is near zero
    std::vector<float> h(np, 0.f);
    uint64_t st = seed * 0x9E3779B97F4A7C15ULL + 4242;
    auto uni = [&] {
      uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      z ^= z >> 31;
      return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](size_t off, int rows, int cols, double gain) {
      const double a = gain * std::sqrt(6.0 / (rows + cols));
      for (size_t i = 0; i < static_cast<size_t>(rows) * cols; i++) h[off + i] = static_cast<float>((2 * uni() - 1) * a);
    };
    fill(oW1, W, FIN, 1.0);
    fill(oW2, W, W, 1.0);
    fill(oV2, W, W, 0.5);
    fill(oW3, W, W, 1.0);
    fill(oV3, W, W, 0.5);
    fill(oW4, FOUT, W, 0.1);
    w.upload(h.data(), np, 0, stream);
    std::vector<int> hd(C), hm(C);
    for (int m = 0; m <= N; m++)
      for (int n = m; n <= N; n++) {
        const int c = tri_off(m, N1) + n - m;
        hd[c] = n;
        hm[c] = m;
      }
    deg.alloc(C);
    ordm.alloc(C);
    deg.upload(hd.data(), C, 0, stream);
    ordm.upload(hm.data(), C, 0, stream);
    std::vector<float> ones(3 * N1, 1.f);
    in_scale.alloc(3 * N1);
    out_scale.alloc(3 * N1);
    in_scale.upload(ones.data(), 3 * N1, 0, stream);
    out_scale.upload(ones.data(), 3 * N1, 0, stream);
    round_weights();
    sync();
  }
  ~Corrector() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  Corrector(const Corrector &) = delete;
  Corrector &operator=(const Corrector &) = delete;

  void sync() { LMW_HIP(hipStreamSynchronize(stream)); }
  void round_weights() {
    hipLaunchKernelGGL(k_round_bf16, dim3(blocks(np)), dim3(256), 0, stream, w.p, wb.p, np);
    LMW_LAUNCH_CHECK();
  }

  void ensure(int S) {
    if (S <= cap) return;
    cap = S;
    const size_t R = static_cast<size_t>(S) * C, RW = R * W, PW = static_cast<size_t>(S) * N1 * W;
    X.alloc(R * FIN);
    H1.alloc(RW);
    H2.alloc(RW);
    H3.alloc(RW);
    Y.alloc(R * FOUT);
    T.alloc(R * FOUT);
    dY.alloc(R * FOUT);
    dA.alloc(RW);
    dB.alloc(RW);
    P1.alloc(PW);
    P2.alloc(PW);
    PM.alloc(PW);
    Q.alloc(PW);
    Gm.alloc(PW);
    const size_t rmax = std::max(R, static_cast<size_t>(S) * N1);
    part.alloc(((rmax + WROWS - 1) / WROWS) * (static_cast<size_t>(W) * W + W));
    lossrow.alloc(R);
    raw.alloc(R * 6);
    raw2.alloc(R * 6);
  }

  // ---------------------------------------------------------------- network
  // dense layers with inputs in multiples of 32 run on the BF16 MFMA, others (the 8 input features) on FMA
  void dense(const float *Xin, int R, int I, size_t oW, const float *bias, int O, const float *mix, int act, float *Yout) {
    const float *Wl = wb.p + oW;
    if (I % 32 == 0) {
      const dim3 g(blocks(R, 64));
      switch ((O + 15) / 16) {
        case 1: hipLaunchKernelGGL(k_dense_fwd_mx<1>, g, dim3(256), 0, stream, Xin, R, I, Wl, bias, O, mix, deg.p, C, N1, act, Yout); break;
        case 2: hipLaunchKernelGGL(k_dense_fwd_mx<2>, g, dim3(256), 0, stream, Xin, R, I, Wl, bias, O, mix, deg.p, C, N1, act, Yout); break;
        case 3: hipLaunchKernelGGL(k_dense_fwd_mx<3>, g, dim3(256), 0, stream, Xin, R, I, Wl, bias, O, mix, deg.p, C, N1, act, Yout); break;
        default: hipLaunchKernelGGL(k_dense_fwd_mx<4>, g, dim3(256), 0, stream, Xin, R, I, Wl, bias, O, mix, deg.p, C, N1, act, Yout); break;
      }
    } else {
      hipLaunchKernelGGL(k_dense_fwd, dim3(blocks(R, 64)), dim3(256), 0, stream, Xin, R, I, Wl, bias, O, mix, deg.p, C, N1, act, Yout);
    }
    LMW_LAUNCH_CHECK();
  }
  // the activations of one forward pass of S samples: the handle's own buffers (ensure), or carved out of a
  // caller-free workspace (forward_workspace)
  struct Acts {
    float *X, *H1, *H2, *H3, *Y, *P1, *P2, *PM;
  };
  Acts own() { return {X.p, H1.p, H2.p, H3.p, Y.p, P1.p, P2.p, PM.p}; }
  // bytes of workspace the activations of a forward pass of S samples need, each array aligned to 256 bytes; with
  // base != nullptr, the arrays' addresses inside [base, base + bytes)
  size_t forward_workspace(int S, unsigned char *base, Acts *a) const {
    const size_t R = static_cast<size_t>(S) * C, RW = R * W, PW = static_cast<size_t>(S) * N1 * W;
    size_t o = 0;
    auto take = [&](size_t bytes) {
      unsigned char *p = base ? base + o : nullptr;
      o += (bytes + 255) / 256 * 256;
      return p;
    };
    float **dst[8] = {&a->X, &a->H1, &a->H2, &a->H3, &a->Y, &a->P1, &a->P2, &a->PM};
    const size_t count[8] = {R * FIN, RW, RW, RW, R * FOUT, PW, PW, PW};
    for (int i = 0; i < 8; i++) *dst[i] = reinterpret_cast<float *>(take(count[i] * sizeof(float)));
    return o;
  }
  // P = mean_m(H), PMo = P V^T (per sample and degree)
  void mixing(const float *H, int S, size_t oV, float *P, float *PMo) {
    hipLaunchKernelGGL(k_pool, dim3(blocks(static_cast<size_t>(S) * N1 * W)), dim3(256), 0, stream, H, S, C, W, N1, P);
    LMW_LAUNCH_CHECK();
    dense(P, S * N1, W, oV, nullptr, W, nullptr, 0, PMo);
  }
  void forward(int S) { forward(S, own()); }
  void forward(int S, const Acts &a) {
    const int R = S * C;
    dense(a.X, R, FIN, oW1, w.p + ob1, W, nullptr, 1, a.H1);
    mixing(a.H1, S, oV2, a.P1, a.PM);
    dense(a.H1, R, W, oW2, w.p + ob2, W, a.PM, 1, a.H2);
    mixing(a.H2, S, oV3, a.P2, a.PM);
    dense(a.H2, R, W, oW3, w.p + ob3, W, a.PM, 1, a.H3);
    dense(a.H3, R, W, oW4, w.p + ob4, FOUT, nullptr, 0, a.Y);
  }

  // g[oW..] = dZ^T X and g[ob..] = column sums of dZ (ob == SIZE_MAX: no bias)
  void wgrad(const float *dZ, const float *Xin, int R, int O, int I, size_t oW, size_t ob) {
    const int nblk = (R + WROWS - 1) / WROWS;
    if (O % 16 == 0 && I % 16 == 0 && (O / 16) * (I / 16) <= 16)
      hipLaunchKernelGGL(k_dense_bwd_weight_mx, dim3(nblk), dim3(256), 0, stream, dZ, Xin, R, O, I, part.p);
    else
      hipLaunchKernelGGL(k_dense_bwd_weight, dim3(nblk), dim3(256), 0, stream, dZ, Xin, R, O, I, part.p);
    LMW_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_parts, dim3(blocks(static_cast<size_t>(O) * I + O)), dim3(256), 0, stream, part.p, nblk, O, I, g.p + oW,
                       ob == SIZE_MAX ? nullptr : g.p + ob);
    LMW_LAUNCH_CHECK();
  }
  void bwd_data(const float *dZ, int R, int O, size_t oW, int I, float *dX) {
    const float *Wl = wb.p + oW;
    if (I % 16 == 0) {
      const dim3 g(blocks(R, 64));
      switch (I / 16) {
        case 1: hipLaunchKernelGGL(k_dense_bwd_data_mx<1>, g, dim3(256), 0, stream, dZ, R, O, Wl, I, dX); break;
        case 2: hipLaunchKernelGGL(k_dense_bwd_data_mx<2>, g, dim3(256), 0, stream, dZ, R, O, Wl, I, dX); break;
        case 3: hipLaunchKernelGGL(k_dense_bwd_data_mx<3>, g, dim3(256), 0, stream, dZ, R, O, Wl, I, dX); break;
        default: hipLaunchKernelGGL(k_dense_bwd_data_mx<4>, g, dim3(256), 0, stream, dZ, R, O, Wl, I, dX); break;
      }
    } else {
      hipLaunchKernelGGL(k_dense_bwd_data, dim3(blocks(R, 64)), dim3(256), 0, stream, dZ, R, O, Wl, I, dX);
    }
    LMW_LAUNCH_CHECK();
  }
  // backward of an inner layer: dZ (holding dH on entry) -> weight, bias and mixing gradients, and
  // dX = dL/d(layer input)
  void inner_backward(int S, const float *Hout, float *dZ, const float *Hin, const float *Pin, size_t oW, size_t ob, size_t oV, float *dX) {
    const int R = S * C;
    const size_t nR = static_cast<size_t>(R) * W, nP = static_cast<size_t>(S) * N1 * W;
    hipLaunchKernelGGL(k_tanh_bwd, dim3(blocks(nR)), dim3(256), 0, stream, Hout, dZ, nR);
    LMW_LAUNCH_CHECK();
    wgrad(dZ, Hin, R, W, W, oW, ob);
    hipLaunchKernelGGL(k_degsum, dim3(blocks(nP)), dim3(256), 0, stream, dZ, S, C, W, N1, Q.p);
    LMW_LAUNCH_CHECK();
    wgrad(Q.p, Pin, S * N1, W, W, oV, SIZE_MAX);
    bwd_data(dZ, R, W, oW, W, dX);
    bwd_data(Q.p, S * N1, W, oV, W, Gm.p);
    hipLaunchKernelGGL(k_unpool_add, dim3(blocks(nR)), dim3(256), 0, stream, Gm.p, S, C, W, N1, deg.p, dX);
    LMW_LAUNCH_CHECK();
  }
  // gradients of the loss (dY filled by k_loss) into g
  void backward(int S) {
    const int R = S * C;
    wgrad(dY.p, H3.p, R, FOUT, W, oW4, ob4);
    bwd_data(dY.p, R, FOUT, oW4, W, dA.p);                           // dA = dH3
    inner_backward(S, H3.p, dA.p, H2.p, P2.p, oW3, ob3, oV3, dB.p);  // dB = dH2
    inner_backward(S, H2.p, dB.p, H1.p, P1.p, oW2, ob2, oV2, dA.p);  // dA = dH1
    const size_t nR = static_cast<size_t>(R) * W;
    hipLaunchKernelGGL(k_tanh_bwd, dim3(blocks(nR)), dim3(256), 0, stream, H1.p, dA.p, nR);
    LMW_LAUNCH_CHECK();
    wgrad(dA.p, X.p, R, W, FIN, oW1, ob1);
  }
  double trained_count(int S) const { return static_cast<double>(S) * (static_cast<double>(C - 1) * FOUT - 3.0 * N); }
  void need_sums(size_t n) {
    if (sums.n < n) sums.alloc(n);
  }
  // loss of S samples (features in X, targets in T) summed into sums[slot]; gradients into g when `grad`
  void loss(int S, bool grad, int slot) {
    const int R = S * C;
    forward(S);
    hipLaunchKernelGGL(k_loss, dim3(blocks(R)), dim3(256), 0, stream, Y.p, T.p, S, C, ordm.p, static_cast<float>(1.0 / trained_count(S)),
                       grad ? dY.p : nullptr, lossrow.p);
    LMW_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, stream, lossrow.p, static_cast<size_t>(R), sums.p, slot);
    LMW_LAUNCH_CHECK();
    if (grad) backward(S);
  }
  void adam_step(float lr) {
    adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, static_cast<float>(adam_t)), bc2 = 1.f - std::pow(b2, static_cast<float>(adam_t));
    hipLaunchKernelGGL(k_adam, dim3(blocks(np)), dim3(256), 0, stream, w.p, wb.p, g.p, m1.p, m2.p, np, lr, b1, b2, 1e-8f, bc1, bc2);
    LMW_LAUNCH_CHECK();
  }
  void features(const float *rawp, int S) { features(rawp, S, X.p); }
  void features(const float *rawp, int S, float *Xo) {
    hipLaunchKernelGGL(k_features, dim3(blocks(static_cast<size_t>(S) * C)), dim3(256), 0, stream, rawp, S, C, N, deg.p, ordm.p, in_scale.p, FIN, Xo);
    LMW_LAUNCH_CHECK();
  }
  void targets(const float *rawp, int S) {
    hipLaunchKernelGGL(k_features, dim3(blocks(static_cast<size_t>(S) * C)), dim3(256), 0, stream, rawp, S, C, N, deg.p, ordm.p, out_scale.p, FOUT,
                       T.p);
    LMW_LAUNCH_CHECK();
  }
  // correction (raw units) of the S states in raw, into raw2; residual[s] = rms of the trained network outputs
  void correction(int S, double *residual) {
    if (residual) need_sums(S);
    correction(S, raw.p, raw2.p, own(), residual ? sums.p : nullptr);
    if (residual) {
      LMW_HIP(hipMemcpyAsync(residual, sums.p, S * sizeof(double), hipMemcpyDeviceToHost, stream));
      sync();
    }
  }
  // the same on device arrays: states `in`, correction `out`, activations `a`; the per-sample rms into the device
  // array `rms` when it is not null.  Launches only (on the handle's stream).
  void correction(int S, const float *in, float *out, const Acts &a, double *rms) {
    features(in, S, a.X);
    forward(S, a);
    hipLaunchKernelGGL(k_unnormalize, dim3(blocks(static_cast<size_t>(S) * C)), dim3(256), 0, stream, a.Y, S, C, N, deg.p, ordm.p, out_scale.p,
                       out);
    LMW_LAUNCH_CHECK();
    if (rms) {
      hipLaunchKernelGGL(k_member_sq, dim3(S), dim3(256), 0, stream, a.Y, C, ordm.p, trained_count(1), rms);
      LMW_LAUNCH_CHECK();
    }
  }
};

// One corrected coarse step of every member of `core` (the corrector's truncation): `nc` RK4 steps of dt/nc, then
// the network's correction.  Members whose residual exceeds `threshold` (or is not finite) are instead stepped from
// the start state with `nf` plain RK4 steps of dt/nf.
void corrected_step(SphereCore &core, Corrector &k, double dt, int nc, int nf, double threshold, int *used_fallback, double *residual) {
  if (core.N != k.N) invalid("the corrector was built for another truncation");
  if (!(dt > 0) || nc < 1 || nf < 1) invalid("need dt > 0 and positive sub-step counts");
  const int B = core.B, S = core.S;
  const size_t nstate = static_cast<size_t>(B) * 3 * S;
  DevArray<double2> save;
  save.alloc(nstate);
  LMW_HIP(hipMemcpyAsync(save.p, core.x0.p, nstate * sizeof(double2), hipMemcpyDeviceToDevice, core.stream));
  core.steps(dt / nc, nc);
  k.ensure(B);
  hipLaunchKernelGGL(k_truncate_out, dim3(blocks(nstate)), dim3(256), 0, core.stream, core.geo(), core.x0.p, core.N, k.raw.p);
  LMW_LAUNCH_CHECK();
  core.sync();
  std::vector<double> res(B);
  k.correction(B, res.data());
  std::vector<int> use(B);
  int nfall = 0;
  for (int b = 0; b < B; b++) {
    use[b] = (res[b] <= threshold) ? 1 : 0;
    nfall += 1 - use[b];
    if (used_fallback) used_fallback[b] = 1 - use[b];
    if (residual) residual[b] = res[b];
  }
  DevArray<int> duse;
  duse.alloc(B);
  duse.upload(use.data(), B, 0, core.stream);
  hipLaunchKernelGGL(k_add_correction, dim3(blocks(nstate)), dim3(256), 0, core.stream, core.geo(), core.x0.p, core.xs.p, k.raw2.p, duse.p);
  LMW_LAUNCH_CHECK();
  if (nfall > 0) {
    // the corrected members are kept aside, every member restarts from the saved state with the plain sub-steps,
    // and the corrected ones are put back
    DevArray<double2> kept;
    kept.alloc(nstate);
    LMW_HIP(hipMemcpyAsync(kept.p, core.x0.p, nstate * sizeof(double2), hipMemcpyDeviceToDevice, core.stream));
    LMW_HIP(hipMemcpyAsync(core.x0.p, save.p, nstate * sizeof(double2), hipMemcpyDeviceToDevice, core.stream));
    LMW_HIP(hipMemcpyAsync(core.xs.p, save.p, nstate * sizeof(double2), hipMemcpyDeviceToDevice, core.stream));
    core.steps(dt / nf, nf);
    hipLaunchKernelGGL(k_restore_members, dim3(blocks(nstate)), dim3(256), 0, core.stream, core.geo(), kept.p, duse.p, core.x0.p, core.xs.p);
    LMW_LAUNCH_CHECK();
    core.sync();
  }
  core.sync();
  for (int b = 0; b < B; b++) core.t[b] += dt;
}

void truncated_out(SphereCore &c, int T, float *out, size_t n) {
  if (T < 1 || T > c.N) invalid("truncation out of range");
  const size_t CT = static_cast<size_t>(T + 1) * (T + 2) / 2, need = static_cast<size_t>(c.B) * 3 * CT * 2;
  if (n != need) invalid("buffer size mismatch");
  DevArray<float> d;
  d.alloc(need);
  hipLaunchKernelGGL(k_truncate_out, dim3(blocks(need / 2)), dim3(256), 0, c.stream, c.geo(), c.x0.p, T, d.p);
  LMW_LAUNCH_CHECK();
  d.download(out, need, 0, c.stream);
  c.sync();
}
void truncated_in(SphereCore &c, int T, const float *in, size_t n) {
  if (T < 1 || T > c.N) invalid("truncation out of range");
  const size_t CT = static_cast<size_t>(T + 1) * (T + 2) / 2, need = static_cast<size_t>(c.B) * 3 * CT * 2;
  if (n != need) invalid("buffer size mismatch");
  DevArray<float> d;
  d.alloc(need);
  d.upload(in, need, 0, c.stream);
  hipLaunchKernelGGL(k_truncate_in, dim3(blocks(static_cast<size_t>(c.B) * 3 * c.S)), dim3(256), 0, c.stream, c.geo(), d.p, T, c.x0.p, c.xs.p);
  LMW_LAUNCH_CHECK();
  c.sync();
}

// ---------------------------------------------------------------- the allocation-free forward pass
// Workspace of lmw_corrector_forward_noalloc: the activations of one forward pass, in a static device array that
// each GPU holds until the process ends (its own copy, at its own address).  Host code never uses the array's name
// as an address: the per-device address comes from hipGetSymbolAddress, taken once per device with that device
// current.  Every kernel of the pass writes each workspace element before reading it, so nothing is cleared between
// calls; two calls that overlap on one GPU share the array and race.
constexpr size_t kNoAllocBytes = static_cast<size_t>(32) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
unsigned char *scratch_address[kMaxDevices] = {};

unsigned char *scratch_on_current_device() {
  int dev = 0;
  LMW_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error(LMW_EDEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_address[dev]) {
    void *p = nullptr;
    LMW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    scratch_address[dev] = static_cast<unsigned char *>(p);
  }
  return scratch_address[dev];
}

// true when p is device (or managed) memory; a query that fails leaves no error behind for later launch checks
bool device_pointer(const void *p) {
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}
}  // namespace

}  // namespace lmw

using namespace lmw;


struct lmw_corrector {
  std::unique_ptr<Corrector> c;
};

#define NEED(p) \
  if (!(p)) invalid("null pointer argument")

extern "C" {

lmw_status lmw_corrector_create(int truncation, int width, uint64_t seed, lmw_corrector **out) {
  return guard([&] {
    NEED(out);
    *out = nullptr;
    auto h = std::make_unique<lmw_corrector>();
    h->c = std::make_unique<Corrector>(truncation, width, seed);
    *out = h.release();
  });
}
lmw_status lmw_corrector_destroy(lmw_corrector *c) {
  return guard([&] { delete c; });
}
lmw_status lmw_corrector_dims(const lmw_corrector *c, int *truncation, int *width, size_t *nparams) {
  return guard([&] {
    NEED(c);
    if (truncation) *truncation = c->c->N;
    if (width) *width = c->c->W;
    if (nparams) *nparams = c->c->np;
  });
}
lmw_status lmw_corrector_get_params(lmw_corrector *c, float *out, size_t n) {
  return guard([&] {
    NEED(c && out);
    if (n != c->c->np) invalid("parameter count mismatch");
    c->c->w.download(out, n, 0, c->c->stream);
    c->c->sync();
  });
}
lmw_status lmw_corrector_set_params(lmw_corrector *c, const float *in, size_t n) {
  return guard([&] {
    NEED(c && in);
    Corrector &k = *c->c;
    if (n != k.np) invalid("parameter count mismatch");
    k.w.upload(in, n, 0, k.stream);
    k.round_weights();
    LMW_HIP(hipMemsetAsync(k.m1.p, 0, n * sizeof(float), k.stream));
    LMW_HIP(hipMemsetAsync(k.m2.p, 0, n * sizeof(float), k.stream));
    k.adam_t = 0;
    k.sync();
  });
}
lmw_status lmw_corrector_set_scales(lmw_corrector *c, const float *in_scale, const float *out_scale, size_t n) {
  return guard([&] {
    NEED(c && in_scale && out_scale);
    Corrector &k = *c->c;
    if (n != static_cast<size_t>(3) * k.N1) invalid("scale arrays hold 3 (N+1) entries");
    for (size_t i = 0; i < n; i++)
      if (!(in_scale[i] > 0) || !(out_scale[i] > 0)) invalid("scales must be positive");
    k.in_scale.upload(in_scale, n, 0, k.stream);
    k.out_scale.upload(out_scale, n, 0, k.stream);
    k.sync();
  });
}
lmw_status lmw_corrector_loss_grad(lmw_corrector *c, const float *x, const float *target, int nsamples, double *loss, float *grad) {
  return guard([&] {
    NEED(c && x && target && loss);
    Corrector &k = *c->c;
    if (nsamples < 1) invalid("need at least one sample");
    k.ensure(nsamples);
    k.need_sums(1);
    const size_t n = static_cast<size_t>(nsamples) * 3 * k.C * 2;
    k.raw.upload(x, n, 0, k.stream);
    k.raw2.upload(target, n, 0, k.stream);
    k.features(k.raw.p, nsamples);
    k.targets(k.raw2.p, nsamples);
    k.loss(nsamples, grad != nullptr, 0);
    double s = 0;
    LMW_HIP(hipMemcpyAsync(&s, k.sums.p, sizeof(double), hipMemcpyDeviceToHost, k.stream));
    if (grad) k.g.download(grad, k.np, 0, k.stream);
    k.sync();
    *loss = s / k.trained_count(nsamples);
  });
}
lmw_status lmw_corrector_forward(lmw_corrector *c, const float *x, int nsamples, float *correction, double *residual) {
  return guard([&] {
    NEED(c && x && correction);
    Corrector &k = *c->c;
    if (nsamples < 1) invalid("need at least one sample");
    k.ensure(nsamples);
    const size_t n = static_cast<size_t>(nsamples) * 3 * k.C * 2;
    k.raw.upload(x, n, 0, k.stream);
    k.correction(nsamples, residual);
    k.raw2.download(correction, n, 0, k.stream);
    k.sync();
  });
}
lmw_status lmw_corrector_set_dataset(lmw_corrector *c, const float *x, const float *target, int nsamples) {
  return guard([&] {
    NEED(c && x && target);
    if (nsamples < 1) invalid("need at least one sample");
    Corrector &k = *c->c;
    const size_t n = static_cast<size_t>(nsamples) * 3 * k.C * 2;
    k.xs.alloc(n);
    k.ts.alloc(n);
    k.perm.alloc(nsamples);
    k.xs.upload(x, n, 0, k.stream);
    k.ts.upload(target, n, 0, k.stream);
    k.nsamples = nsamples;
    k.sync();
  });
}
lmw_status lmw_corrector_train(lmw_corrector *c, int epochs, int batch, float lr, uint64_t seed, double *losses) {
  return guard([&] {
    NEED(c && losses);
    Corrector &k = *c->c;
    if (k.nsamples < 1) invalid("no training set");
    if (epochs < 0 || batch < 1 || !(lr > 0)) invalid("need epochs >= 0, batch >= 1 and lr > 0");
    batch = std::min(batch, k.nsamples);
    k.ensure(batch);
    const int nb = k.nsamples / batch;
    k.need_sums(nb);
    const size_t per = static_cast<size_t>(3) * k.C * 2;
    std::vector<int> order(k.nsamples);
    for (int i = 0; i < k.nsamples; i++) order[i] = i;
    uint64_t st = seed * 0x9E3779B97F4A7C15ULL + 99;
    auto next = [&] {
      uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      return z ^ (z >> 31);
    };
    std::vector<double> bl(nb);
    for (int e = 0; e < epochs; e++) {
      for (int i = k.nsamples - 1; i > 0; i--) std::swap(order[i], order[next() % static_cast<uint64_t>(i + 1)]);
      k.perm.upload(order.data(), k.nsamples, 0, k.stream);
      for (int q = 0; q < nb; q++) {
        const int *idx = k.perm.p + static_cast<size_t>(q) * batch;
        hipLaunchKernelGGL(k_gather, dim3(blocks(per * batch)), dim3(256), 0, k.stream, k.xs.p, idx, batch, per, k.raw.p);
        LMW_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_gather, dim3(blocks(per * batch)), dim3(256), 0, k.stream, k.ts.p, idx, batch, per, k.raw2.p);
        LMW_LAUNCH_CHECK();
        k.features(k.raw.p, batch);
        k.targets(k.raw2.p, batch);
        k.loss(batch, true, q);
        k.adam_step(lr);
      }
      LMW_HIP(hipMemcpyAsync(bl.data(), k.sums.p, nb * sizeof(double), hipMemcpyDeviceToHost, k.stream));
      k.sync();  // also keeps `order` unchanged until its upload has completed
      double s = 0;
      for (double v : bl) s += v;
      losses[e] = s / (k.trained_count(batch) * nb);
    }
    k.sync();
  });
}
lmw_status lmw_sphere_corrected_step(lmw_sphere *s, lmw_corrector *c, double dt, int coarse_substeps, int fallback_substeps, double threshold,
                                     int *used_fallback, double *residual) {
  return guard([&] {
    NEED(s && c);
    corrected_step(core_of(s), *c->c, dt, coarse_substeps, fallback_substeps, threshold, used_fallback, residual);
  });
}
lmw_status lmw_ensemble_corrected_step(lmw_ensemble *e, lmw_corrector *c, double dt, int coarse_substeps, int fallback_substeps, double threshold,
                                       int *used_fallback, double *residual) {
  return guard([&] {
    NEED(e && c);
    corrected_step(single_group(e), *c->c, dt, coarse_substeps, fallback_substeps, threshold, used_fallback, residual);
  });
}
lmw_status lmw_ensemble_get_truncated(lmw_ensemble *e, int truncation, float *out, size_t n) {
  return guard([&] {
    NEED(e && out);
    truncated_out(single_group(e), truncation, out, n);
  });
}
lmw_status lmw_ensemble_set_truncated(lmw_ensemble *e, int truncation, const float *in, size_t n) {
  return guard([&] {
    NEED(e && in);
    truncated_in(single_group(e), truncation, in, n);
  });
}

lmw_status lmw_corrector_forward_noalloc(lmw_corrector *c, const float *x, int nsamples, float *correction, double *residual) {
  return guard([&] {
    NEED(c && x && correction);
    Corrector &k = *c->c;
    if (nsamples < 1) invalid("need at least one sample");
    if (!device_pointer(x) || !device_pointer(correction) || (residual && !device_pointer(residual)))
      invalid("lmw_corrector_forward_noalloc: inputs and outputs must be device memory");
    Corrector::Acts a{};
    const size_t need = k.forward_workspace(nsamples, nullptr, &a);
    if (need > kNoAllocBytes)
      invalid("lmw_corrector_forward_noalloc: " + std::to_string(nsamples) + " samples need " + std::to_string(need) +
              " bytes of workspace, more than the " + std::to_string(kNoAllocBytes) + " bytes it has");
    k.forward_workspace(nsamples, scratch_on_current_device(), &a);
    k.correction(nsamples, x, correction, a, residual);
    k.sync();
  });
}

}  // extern "C"
