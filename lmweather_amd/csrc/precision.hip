// Mixed-precision Legendre transforms on MFMA for gfx950, FP32 accumulation throughout:
//   LMW_FP32    FP32 operands (v_mfma_f32_16x16x4_f32);
//   LMW_BF16X3  operands split in two BF16 parts x = x1 + x2 (16 significant bits), three products
//               x1 y1 + x1 y2 + x2 y1 (v_mfma_f32_16x16x32_bf16): relative error ~1e-5 per product;
//   LMW_BF16X6  operands split in three BF16 parts (24 bits, an FP32 value), six products x1 y1 + x1 y2 + x2 y1 +
//               x1 y3 + x2 y2 + x3 y1 (the dropped terms are below 2^-24 relative): FP32-level error at twice the
//               MFMA work of BF16X3, which the memory-bound kernels mostly hide.
// The Legendre tables are built in FP64 (k_tables) and rounded/split once; the state, the FFTs, the grid products
// and the RK4 update stay FP64.
//
// Both transforms are, per zonal wavenumber m and parity of n - m, small GEMMs whose columns are ensemble members:
//   inverse  F[jh][b] (quantity q) = sum_k' A[jh][k'] C_q[k'][b]      A = P or H rows of parity p, C = coefficients
//   forward  O_q[r][b]            = sum_jh P[jh][r] X_q[jh][b] + H[jh][r] Y_q[jh][b]
// Each 16 x 16 MFMA tile holds one quantity for 16 members, so one lane's accumulators carry every quantity of
// its member and the epilogues (north/south combination, RK4 update) need no data exchange.
#include <algorithm>
#include <map>
#include <cmath>
#include <vector>

#include "sphere_core.h"
#include "sphere_epilogue.h"

namespace lmw {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace {
constexpr int MEM = 16;  // members per block (the N dimension of an MFMA tile)
// the inverse transform's northern latitudes are split over this many blocks per (m, member tile): more, smaller
// blocks, so that one block's coefficient staging overlaps another's MFMAs
constexpr int kSynthJSplit = 2;

// The remainders are formed from the rounded FP32 value x, never fused with the product that produced x (FMA
// contraction would split the exact product instead, and the bits would then depend on where the split is inlined).
__device__ inline void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
#pragma clang fp contract(off)
  hi = static_cast<__bf16>(x);
  lo = static_cast<__bf16>(x - static_cast<float>(hi));
}
// x = x1 + x2 + x3 to the last bit of an FP32 value (each remainder is exact in FP32)
__device__ inline void split_bf16x3(float x, __bf16 &x1, __bf16 &x2, __bf16 &x3) {
#pragma clang fp contract(off)
  x1 = static_cast<__bf16>(x);
  const float r1 = x - static_cast<float>(x1);
  x2 = static_cast<__bf16>(r1);
  x3 = static_cast<__bf16>(r1 - static_cast<float>(x2));
}
constexpr int nsplit(int prec) { return prec == LMW_BF16X6 ? 3 : 2; }

typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

// Member column swizzle of the FP64 forward transform's operand chunk (row q, k): distinct for the 16 (q / 2, k / 2 mod 4)
__device__ inline int f64_swz(int q, int k) { return (((q >> 1) << 2) + ((k >> 1) & 3)) & 15; }

// Waits until this wave's vector memory loads are done, loads into LDS included (s_waitcnt vmcnt(0))
__device__ inline void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Block barrier for LDS data only: this wave's LDS accesses complete, then s_barrier; the "memory" clobbers keep
// the compiler from moving memory accesses across it.  Unlike __syncthreads() it does not wait for the wave's
// outstanding global loads, among them loads into LDS (global_load_lds) whose buffer only the issuing wave reads.
__device__ inline void lds_barrier() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_s_barrier();
  __asm__ volatile("" ::: "memory");
}
// split_bf16x3 of two values at once, each part of the pair packed into one dword (a in the low half): one packed
// conversion per part for both values, and the part's FP32 value recovered from the packed bits by a shift or a mask
// (the same round-to-nearest-even parts as split_bf16x3; 11 instructions per pair instead of 15)
__device__ inline void split3_pair(float a, float b, unsigned &p1, unsigned &p2, unsigned &p3) {
#pragma clang fp contract(off)
  p1 = __builtin_bit_cast(unsigned, __builtin_convertvector(v2f{a, b}, v2bf));
  const float ra = a - __builtin_bit_cast(float, p1 << 16), rb = b - __builtin_bit_cast(float, p1 & 0xffff0000u);
  p2 = __builtin_bit_cast(unsigned, __builtin_convertvector(v2f{ra, rb}, v2bf));
  const float sa = ra - __builtin_bit_cast(float, p2 << 16), sb = rb - __builtin_bit_cast(float, p2 & 0xffff0000u);
  p3 = __builtin_bit_cast(unsigned, __builtin_convertvector(v2f{sa, sb}, v2bf));
}
// split_bf16 (two parts) of two values at once, packed as in split3_pair
__device__ inline void split2_pair(float a, float b, unsigned &p1, unsigned &p2) {
#pragma clang fp contract(off)
  p1 = __builtin_bit_cast(unsigned, __builtin_convertvector(v2f{a, b}, v2bf));
  const float ra = a - __builtin_bit_cast(float, p1 << 16), rb = b - __builtin_bit_cast(float, p1 & 0xffff0000u);
  p2 = __builtin_bit_cast(unsigned, __builtin_convertvector(v2f{ra, rb}, v2bf));
}
// eight FP32 values (lo = 0..3, hi = 4..7) split into their three BF16 parts as MFMA fragments
__device__ inline void split8_regs(v4f lo, v4f hi, v8bf &a1, v8bf &a2, v8bf &a3) {
  unsigned x[3][4];
  split3_pair(lo[0], lo[1], x[0][0], x[1][0], x[2][0]);
  split3_pair(lo[2], lo[3], x[0][1], x[1][1], x[2][1]);
  split3_pair(hi[0], hi[1], x[0][2], x[1][2], x[2][2]);
  split3_pair(hi[2], hi[3], x[0][3], x[1][3], x[2][3]);
  const v4u u1{x[0][0], x[0][1], x[0][2], x[0][3]}, u2{x[1][0], x[1][1], x[1][2], x[1][3]}, u3{x[2][0], x[2][1], x[2][2], x[2][3]};
  a1 = __builtin_bit_cast(v8bf, u1);
  a2 = __builtin_bit_cast(v8bf, u2);
  a3 = __builtin_bit_cast(v8bf, u3);
}

// acc += a * b for one 16x16 tile; PREC = LMW_FP32 consumes one float per lane, LMW_BF16X3 eight bf16 pairs
__device__ inline v4f mma32(float a, float b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
__device__ inline v4d mma64(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline v4f mma3(v8bf ah, v8bf al, v8bf bh, v8bf bl, v4f c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
  return c;
}
// the six products of three-part operands, smallest first
__device__ inline v4f mma6(v8bf a1, v8bf a2, v8bf a3, v8bf b1, v8bf b2, v8bf b3, v4f c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, b1, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b2, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b3, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b1, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b2, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, c, 0, 0, 0);
  return c;
}
__device__ inline v8bf load8(const __bf16 *p) { return *reinterpret_cast<const v8bf *>(p); }
// eight FP32 table entries (32-byte aligned) split into their three BF16 parts, as the tables are split for the
// operands: 4 bytes per entry from memory instead of the 6 of three stored parts
__device__ inline void split8(const float *p, v8bf &a1, v8bf &a2, v8bf &a3) {
  split8_regs(*reinterpret_cast<const v4f *>(p), *reinterpret_cast<const v4f *>(p + 4), a1, a2, a3);
}
}  // namespace

// ------------------------------------------------------------------------------------------- table builders

// Forward-transform A operand: FP32 [m][k = 2 jh (P) | 2 jh + 1 (H)][r], split BF16 parts 1 and 2 [m][r][k] (BF16X3),
// and FP32 [m][r][k] in that order (Abf: BF16X6 splits it into three parts as it loads), from the FP64 analysis
// tables Pa/Ha [m][jh][r].
__global__ void k_build_anal_tables(Geo G, int R, const double *__restrict__ Pa, const double *__restrict__ Ha, float *__restrict__ A32,
                                    __bf16 *__restrict__ Abh, __bf16 *__restrict__ Abl, float *__restrict__ Abf, const int *__restrict__ mrow) {
  const size_t total = static_cast<size_t>(2 * G.Jhp) * R;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    // e enumerates (global row gr, k) with k fastest: gr = poff[m] + r
    const int gr = static_cast<int>(e / (2 * G.Jhp)), k = static_cast<int>(e % (2 * G.Jhp));
    const int m = mrow[gr], r = gr - G.poff[m], kp = G.kpe[m] + G.kpo[m];
    const int jh = k >> 1;  // k = 2 jh + part: the P and H entries of a latitude pair side by side
    const double *T = (k & 1) ? Ha : Pa;
    const float v = static_cast<float>(T[static_cast<size_t>(G.poff[m]) * G.Jhp + static_cast<size_t>(jh) * kp + r]);
    A32[static_cast<size_t>(2 * G.Jhp) * G.poff[m] + static_cast<size_t>(k) * kp + r] = v;
    __bf16 x1, x2, x3;
    split_bf16x3(v, x1, x2, x3);  // x1, x2 are also the two-part split
    Abh[e] = x1;
    Abl[e] = x2;
    Abf[e] = v;
  }
}

// The FP64 forward-transform A operand in A32's layout, [m][k = jh (P) | Jhp + jh (H)][r], one block per m.
__global__ void k_build_anal64(Geo G, const double *__restrict__ Pa, const double *__restrict__ Ha, double *__restrict__ A64) {
  const int m = blockIdx.x, kp = G.kpe[m] + G.kpo[m], KK = 2 * G.Jhp;
  for (int e = threadIdx.x; e < KK * kp; e += blockDim.x) {
    const int k = e / kp, r = e - k * kp, jh = k < G.Jhp ? k : k - G.Jhp;
    const double *T = k < G.Jhp ? Pa : Ha;
    A64[static_cast<size_t>(KK) * G.poff[m] + e] = T[static_cast<size_t>(G.poff[m]) * G.Jhp + static_cast<size_t>(jh) * kp + r];
  }
}

// Inverse-transform A operands: FP32 Ps32/Hs32 in the layout of Ps/Hs ([row][jh]), split BF16 parts 1 and 2
// [m][p][jh][k'] (BF16X3) and FP32 Spf/Shf in that order (BF16X6 splits them as it loads), each parity block
// padded to 32 rows (soff[m] is the offset of m, ke32/ko32 the padded row counts).
__global__ void k_build_synth_tables(Geo G, int R, const double *__restrict__ Ps, const double *__restrict__ Hs, float *__restrict__ Ps32,
                                     float *__restrict__ Hs32, const int *__restrict__ soff, const int *__restrict__ ke32,
                                     const int *__restrict__ ko32, __bf16 *__restrict__ Sph, __bf16 *__restrict__ Spl, __bf16 *__restrict__ Shh,
                                     __bf16 *__restrict__ Shl, float *__restrict__ Spf, float *__restrict__ Shf) {
  const size_t total = static_cast<size_t>(R) * G.Jhp;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    Ps32[e] = static_cast<float>(Ps[e]);
    Hs32[e] = static_cast<float>(Hs[e]);
  }
  // bf16: one thread per (m, p, jh, k')
  for (int m = blockIdx.x; m < G.M1; m += gridDim.x) {
    for (int p = 0; p < 2; p++) {
      const int kp32 = p ? ko32[m] : ke32[m], kp16 = p ? G.kpo[m] : G.kpe[m], rbase = p ? G.kpe[m] : 0;
      const size_t base = static_cast<size_t>(soff[m]) + (p ? static_cast<size_t>(G.Jhp) * ke32[m] : 0);
      for (int t = threadIdx.x; t < G.Jhp * kp32; t += blockDim.x) {
        const int jh = t / kp32, kk = t % kp32;
        float vp = 0, vh = 0;
        if (kk < kp16) {
          const size_t src = (static_cast<size_t>(G.poff[m]) + rbase + kk) * G.Jhp + jh;
          vp = static_cast<float>(Ps[src]);
          vh = static_cast<float>(Hs[src]);
        }
        __bf16 x1, x2, x3;
        split_bf16x3(vp, x1, x2, x3);
        Sph[base + t] = x1;
        Spl[base + t] = x2;
        Spf[base + t] = vp;
        split_bf16x3(vh, x1, x2, x3);
        Shh[base + t] = x1;
        Shl[base + t] = x2;
        Shf[base + t] = vh;
      }
    }
  }
}

// ------------------------------------------------------------------------------------------- inverse transform

struct SynthTabs {
  const float *Ps32, *Hs32;
  const __bf16 *Sph, *Spl, *Shh, *Shl;
  const float *Spf, *Shf;  // FP32 in the BF16 layout, split into three parts on load (BF16X6)
  const int *soff, *ke32, *ko32;
  const double *kn;  // kn[n] = -a^2 / (n (n + 1)) for 1 <= n <= N + 1, kn[0] = 0: psi, chi = kn (zeta, delta)
};

// Block: one m, 16 members, one wave per 16 northern latitudes.  Coefficients are staged in LDS in chunks of 32
// rows of one parity: quantities q = zeta re/im, phi re/im, psi re/im, chi re/im (H uses psi, chi = q 4..7).
// LDS per chunk: FP32 [q][k'][member] floats; BF16 [q][member][k'] per split part.
// Tile shape (autotuned per truncation group): RC rows of one parity per LDS chunk (a multiple of 32, the K of one
// BF16 MFMA).
// The work of one block (order m, members mb .. mb + 15); shared by the per-group and the multi-group kernels.
template <int PREC, int RC>
__device__ __forceinline__ void synth_mx_block(const Geo &G, const SynthTabs &T, const double2 *__restrict__ spec, float2 *__restrict__ Fs,
                                               int b0, int nb, int mixed, int m, int mb, int jt0 = 0) {
  constexpr int SP = RC + 8;  // padded row pitch (bf16) of the BF16 chunk: distinct LDS banks across members
  extern __shared__ float lds_f[];
  float *cf = lds_f;                                                       // FP32 chunk [q][k'][member]
  __bf16 *ch = reinterpret_cast<__bf16 *>(lds_f), *cl = ch + 8 * MEM * SP;  // BF16 chunk [q][member][k'], parts 1 and 2
  __bf16 *c3 = cl + 8 * MEM * SP;                                           // part 3 (BF16X6)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int Km = G.N - m + 1, kpe = G.kpe[m], kpo = G.kpo[m];
  const size_t moff_m = static_cast<size_t>(G.moff[m]);
  const double a = G.a;
  const int jt = jt0 + wave;  // row tile (16 northern latitudes) of this wave
  const bool active = jt * 16 < G.Jh;
  constexpr bool PIPE = (PREC == LMW_BF16X6 || PREC == LMW_BF16X3) && RC == 32;
  v4f accP[2][8], accH[2][4];
  for (int p = 0; p < 2; p++) {
    for (int q = 0; q < 8; q++) accP[p][q] = v4f{0, 0, 0, 0};
    for (int q = 0; q < 4; q++) accH[p][q] = v4f{0, 0, 0, 0};
  }
  for (int p = 0; p < 2; p++) {
    const int kp = p ? kpo : kpe;             // rows of this parity (16-padded)
    const int kp32 = PREC != LMW_FP32 ? (p ? T.ko32[m] : T.ke32[m]) : kp;
    for (int c0 = 0; c0 < kp32; c0 += RC) {
      __syncthreads();
      // split BF16 with 32-row chunks: this chunk's P and H table operands (FP32 entries for BF16X6, their two BF16
      // parts for BF16X3) are requested before the coefficients are staged, so that they have arrived when the MFMAs
      // start
      v4f araw[4];
      if constexpr (PIPE) {
        if (active) {
          const size_t base = static_cast<size_t>(T.soff[m]) + (p ? static_cast<size_t>(G.Jhp) * T.ke32[m] : 0) +
                              static_cast<size_t>(jt * 16 + (lane & 15)) * kp32 + c0 + 8 * (lane >> 4);
          if constexpr (PREC == LMW_BF16X6) {
            araw[0] = *reinterpret_cast<const v4f *>(T.Spf + base);
            araw[1] = *reinterpret_cast<const v4f *>(T.Spf + base + 4);
            araw[2] = *reinterpret_cast<const v4f *>(T.Shf + base);
            araw[3] = *reinterpret_cast<const v4f *>(T.Shf + base + 4);
          } else {
            araw[0] = *reinterpret_cast<const v4f *>(T.Sph + base);
            araw[1] = *reinterpret_cast<const v4f *>(T.Spl + base);
            araw[2] = *reinterpret_cast<const v4f *>(T.Shh + base);
            araw[3] = *reinterpret_cast<const v4f *>(T.Shl + base);
          }
        }
      }
      // stage 8 quantities x RC rows x 16 members: one thread per (member, pair of rows), members fastest, so
      // that a wavefront writes consecutive LDS words (FP32) or whole words of a padded pitch (BF16).  Both rows'
      // coefficients (and their kn) are loaded before either is used, from a clamped valid row and member where the
      // lane has none (zeroed afterwards), so the thread waits for memory once per chunk.
      for (int e = threadIdx.x; e < (RC / 2) * MEM; e += blockDim.x) {
        const int j = e % MEM, kk = 2 * (e / MEM), bl = mb + j;
        const int bc = b0 + min(bl, nb - 1);
        double2 z[2], d[2], f[2];
        double knv[2];
        bool ok[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int kq = c0 + kk + h, k = 2 * kq + p;
          ok[h] = bl < nb && k < Km && kq < kp;
          const size_t q = moff_m + min(k, Km - 1);
          z[h] = spec[G.at(bc, 0, q)];
          d[h] = spec[G.at(bc, 1, q)];
          f[h] = spec[G.at(bc, 2, q)];
          knv[h] = T.kn[min(m + k, G.N + 1)];
        }
        float v[2][8];
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int n = m + 2 * (c0 + kk + h) + p;
          if (!ok[h]) z[h] = d[h] = f[h] = cx(0, 0);
          if (mixed && n == 0) f[h] = cx(0, 0);  // phi_00 P_0^0 (the area mean) stays out of the mixed path
          const double kn = knv[h];  // psi, chi = kn (zeta, delta): sphere.cpp:185-190
          v[h][0] = static_cast<float>(z[h].x);
          v[h][1] = static_cast<float>(z[h].y);
          v[h][2] = static_cast<float>(f[h].x);
          v[h][3] = static_cast<float>(f[h].y);
          v[h][4] = static_cast<float>(z[h].x * kn);
          v[h][5] = static_cast<float>(z[h].y * kn);
          v[h][6] = static_cast<float>(d[h].x * kn);
          v[h][7] = static_cast<float>(d[h].y * kn);
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
          if (PREC == LMW_FP32) {
            cf[(q * RC + kk) * MEM + j] = v[0][q];
            cf[(q * RC + kk + 1) * MEM + j] = v[1][q];
          } else {
            unsigned p1, p2, p3;
            if constexpr (PREC == LMW_BF16X6) split3_pair(v[0][q], v[1][q], p1, p2, p3);
            else split2_pair(v[0][q], v[1][q], p1, p2);
            *reinterpret_cast<unsigned *>(ch + (q * MEM + j) * SP + kk) = p1;
            *reinterpret_cast<unsigned *>(cl + (q * MEM + j) * SP + kk) = p2;
            if constexpr (PREC == LMW_BF16X6) *reinterpret_cast<unsigned *>(c3 + (q * MEM + j) * SP + kk) = p3;
          }
        }
      }
      __syncthreads();
      if (!active) continue;
      const int jh = jt * 16 + (lane & 15);
      if (PREC == LMW_FP32) {
        const int rbase = G.poff[m] + (p ? kpe : 0);
        const int steps = max(0, min(RC, kp - c0)) / 4;
        for (int s = 0; s < steps; s++) {
          const int kk = 4 * s + (lane >> 4);
          const size_t ai = static_cast<size_t>(rbase + c0 + kk) * G.Jhp + jh;
          const float aP = T.Ps32[ai], aH = T.Hs32[ai];
#pragma unroll
          for (int q = 0; q < 8; q++) {
            const float bq = cf[(q * RC + kk) * MEM + (lane & 15)];
            accP[p][q] = mma32(aP, bq, accP[p][q]);
            if (q >= 4) accH[p][q - 4] = mma32(aH, bq, accH[p][q - 4]);
          }
        }
      } else {
        const int nk = min(RC, kp32 - c0) / 32;  // MFMA K-steps of this chunk (kp32 is a multiple of 32)
        for (int s = 0; s < nk; s++) {
          const size_t base = static_cast<size_t>(T.soff[m]) + (p ? static_cast<size_t>(G.Jhp) * T.ke32[m] : 0) +
                              static_cast<size_t>(jh) * kp32 + c0 + 32 * s + 8 * (lane >> 4);
          if constexpr (PREC == LMW_BF16X6) {
            v8bf aPh, aPl, aP3, aHh, aHl, aH3;
            if constexpr (PIPE) {
              split8_regs(araw[0], araw[1], aPh, aPl, aP3);
              split8_regs(araw[2], araw[3], aHh, aHl, aH3);
            } else {
              split8(T.Spf + base, aPh, aPl, aP3);
              split8(T.Shf + base, aHh, aHl, aH3);
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
              const int bo = (q * MEM + (lane & 15)) * SP + 32 * s + 8 * (lane >> 4);
              const v8bf bh = load8(ch + bo), bl = load8(cl + bo), b3 = load8(c3 + bo);
              accP[p][q] = mma6(aPh, aPl, aP3, bh, bl, b3, accP[p][q]);
              if (q >= 4) accH[p][q - 4] = mma6(aHh, aHl, aH3, bh, bl, b3, accH[p][q - 4]);
            }
          } else {
            v8bf aPh, aPl, aHh, aHl;
            if constexpr (PIPE) {
              aPh = __builtin_bit_cast(v8bf, araw[0]);
              aPl = __builtin_bit_cast(v8bf, araw[1]);
              aHh = __builtin_bit_cast(v8bf, araw[2]);
              aHl = __builtin_bit_cast(v8bf, araw[3]);
            } else {
              aPh = load8(T.Sph + base);
              aPl = load8(T.Spl + base);
              aHh = load8(T.Shh + base);
              aHl = load8(T.Shl + base);
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
              const int bo = (q * MEM + (lane & 15)) * SP + 32 * s + 8 * (lane >> 4);
              const v8bf bh = load8(ch + bo), bl = load8(cl + bo);
              accP[p][q] = mma3(aPh, aPl, bh, bl, accP[p][q]);
              if (q >= 4) accH[p][q - 4] = mma3(aHh, aHl, bh, bl, accH[p][q - 4]);
            }
          }
        }
      }
    }
  }
  (void)nwave;
  if (!active) return;
  const int j = lane & 15, bl = mb + j;
  if (bl >= nb) return;
  const int b = b0 + bl;
  const double dm = m, ia = 1.0 / a;  // u, v are rounded to FP32 below: 1/a times instead of a division
  for (int r = 0; r < 4; r++) {
    const int jh = jt * 16 + 4 * (lane >> 4) + r;
    if (jh >= G.Jh) continue;
    for (int s = 0; s < 2; s++) {
      const double sg = s == 0 ? 1.0 : -1.0;
      const int jj = s == 0 ? G.Jh + jh : G.Jh - 1 - jh;
      auto P = [&](int q) { return static_cast<double>(accP[0][q][r]) + sg * static_cast<double>(accP[1][q][r]); };
      auto H = [&](int q) { return sg * static_cast<double>(accH[0][q][r]) + static_cast<double>(accH[1][q][r]); };
      // F = phi' = phi - phibar: the mean phi_00 P_0^0 stays out of the FP32 Fourier coefficients (the grid
      // kernel needs phi' only, and the forward transform's epilogue adds the mean's flux -phibar delta in FP64)
      const double2 Z = cx(P(0), P(1)), F = cx(P(2), P(3)), psP = cx(P(4), P(5)), chP = cx(P(6), P(7));
      const double2 psH = cx(H(0), H(1)), chH = cx(H(2), H(3));
      const double2 u = csub(cim(dm, chP), psH), v = cadd(cim(dm, psP), chH);  // sphere.cpp:200-201
      float2 *o = Fs + ((static_cast<size_t>(b) * G.J + jj) * G.M1 + m) * 4;
      o[0] = from_d2<float2>(cx(u.x * ia, u.y * ia));
      o[1] = from_d2<float2>(cx(v.x * ia, v.y * ia));
      o[2] = from_d2<float2>(Z);
      o[3] = from_d2<float2>(F);
    }
  }
}

// ------------------------------------------------------------------------------------------- forward transform

struct AnalTabs {
  const float *A32;
  const __bf16 *Abh, *Abl;
  const float *Abf;  // FP32 in the BF16 layout, split into three parts on load (BF16X6)
  const double *A64;  // FP64 copy of A32 (the FP64 MFMA variant)
};
// Operand storage of the forward transform: FP64 operands for the FP64 MFMA variant, FP32 for the mixed modes
template <int PREC> using AnalIn = typename std::conditional<PREC == LMW_FP64, double2, float2>::type;

// Block: one (m, parity), 16 members, 4 waves over the 16-row tiles of that parity.  The K dimension runs over the
// northern latitudes twice (P with X, then H with Y), staged in LDS in chunks of 64.  Quantities q = o1 re/im ..
// o4 re/im (dzeta, ddelta, the Laplacian part of ddelta, dphi).
// Tile shape (autotuned per truncation group): KC latitudes of K per LDS chunk, NW waves per block, each wave
// owning row tiles wave, wave + NW, ... (TPW of them: 8 tiles of 16 rows cover the largest parity block, T170).
// The work of one block (order m, parity p, members mb .. mb + 15), shared by the per-group and multi-group kernels.
// PREC = LMW_FP64 runs the same schedule on FP64 operands with the FP64 MFMA (16x16x4): the FP64 step's forward
// transform on matrix cores.  REC_LDS: the split-BF16 records through the LDS buffer described below (off in the
// multi-group kernels, whose register budget it would overrun).
template <int PREC, int KC, int NW, bool REC_LDS = true>
__device__ __forceinline__ void anal_mx_block(const Geo &G, const AnalTabs &T, const AnalIn<PREC> *__restrict__ Gb, int mode, int stage,
                                              double dt, double nu4, double2 *__restrict__ x0, double2 *__restrict__ xs,
                                              double2 *__restrict__ acc, double2 *__restrict__ raw, int b0, int nb, int mixed, int m, int p,
                                              int mb) {
  using ST = typename std::conditional<PREC == LMW_FP64, double, float>::type;  // staged operand type
  using AV = typename std::conditional<PREC == LMW_FP64, v4d, v4f>::type;      // accumulator tile
  using GT = AnalIn<PREC>;
  constexpr int AP = KC + 8;         // padded row pitch (bf16) of the BF16 chunk
  constexpr int TPW = (8 + NW - 1) / NW;
  extern __shared__ float lds_f[];
  ST *bf = reinterpret_cast<ST *>(lds_f);                                    // FP32 / FP64 chunk [q][k][member]
  __bf16 *bh = reinterpret_cast<__bf16 *>(lds_f), *blo = bh + 8 * MEM * AP;  // BF16 chunk [q][member][k], parts 1 and 2
  __bf16 *b3p = blo + 8 * MEM * AP;                                           // part 3 (BF16X6)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kpe = G.kpe[m], kp = kpe + G.kpo[m], Rp = p ? G.kpo[m] : kpe, rbase = p ? kpe : 0;
  if (Rp == 0) return;
  const int nrt = Rp / 16, KK = 2 * G.Jhp;
  const size_t moff_m = static_cast<size_t>(G.moff[m]);  // read once, before the epilogue's stores
  const int poff_m = G.poff[m];                          // and before the K loop's barriers
  AV accv[TPW][8];
  for (int t = 0; t < TPW; t++)
    for (int q = 0; q < 8; q++) accv[t][q] = AV{0, 0, 0, 0};
  // operand staging: one thread per (member, pair of k), members fastest; a thread reads its two records of four
  // contiguous complex operands and writes whole 32-bit LDS words.  The next chunk's records are loaded into
  // registers while the current chunk's MFMAs run.
  constexpr int NIT = ((KC / 2) * MEM + 64 * NW - 1) / (64 * NW);
  ST vv[NIT][2][8];
  constexpr int NP64 = PREC == LMW_FP64 ? (MEM * KC * 4 + 64 * NW - 1) / (64 * NW) : 1;
  GT pv[NP64];  // FP64: the next chunk's operand pieces
  bool pok[NP64];
  float4 rr[NIT][5];  // mixed: the raw records of the next chunk
  bool rr_ok[NIT];
  // mixed: the thread of latitude pair kk / 2 reads that pair's one 80-byte record (five weighted coefficients per
  // latitude) from a clamped, valid address whatever its lane (rec_ok marks the lanes whose record is used), so that
  // every chunk issues the same loads in the same order and the waits for them stay exact
  auto rec_ok = [&](int it, int c0) {
    const int e = threadIdx.x + 64 * NW * it, j = e % MEM, kk = 2 * (e / MEM);
    return e < (KC / 2) * MEM && mb + j < nb && ((c0 + kk) >> 1) < G.Jh;
  };
  auto rec_src = [&](int it, int c0) {
    const int e = threadIdx.x + 64 * NW * it, j = e % MEM, kk = 2 * (e / MEM);
    const int blc = min(mb + j, nb - 1), jhc = min((c0 + kk) >> 1, G.Jh - 1);
    return reinterpret_cast<const float4 *>(Gb + ((static_cast<size_t>(b0 + blc) * G.M1 + m) * G.Jh + jhc) * 10);
  };
  // Split-BF16 modes with 32-latitude chunks and 4 waves (one record per thread): the chunk's records are requested
  // as contiguous 1 KB pieces, one per instruction, straight into an LDS buffer (buffer_load_dwordx4 ... lds: lane
  // i's 16 bytes land at base + 16 i), and each thread then reads its own record from the buffer.  The chunk's
  // records of one member are 16 consecutive 80-byte records (1280 bytes), held at bytes 1280 j .. 1280 j + 1279.
  // Loaded one record per thread instead, a request of 16 bytes per lane touched 16 members' records at once, about
  // 48 cache lines per instruction (T170 x 32: 41.7 -> 36.7 us per launch).  Threads take (member e / 16, pair
  // e % 16) here.  The next chunk's records are requested once every wave has read this chunk's (after the staging
  // barrier), so they arrive during the MFMAs.  Each wave waits for its own requests (vmcnt(0)) before the next
  // chunk's first barrier, since the compiler places no wait before reads of such a buffer, and the barriers wait
  // for LDS only (__syncthreads() would also wait for the requests in flight).  The buffer is an LDS object of its
  // own, apart from the operand chunk, and the requests use the raw-buffer form (not the global one), whose
  // outstanding count the compiler tracks exactly, so the waits for the table operands do not wait for the records.
  // The resource starts at the block's first record (member b0 + mb, order m): every clamped record lies at or
  // after it, within 32-bit offsets.  The FP32 mode keeps the register loads: its table loads in the MFMA phase
  // would wait for these requests (77 -> 87 us).
  constexpr bool DMA = REC_LDS && (PREC == LMW_BF16X6 || PREC == LMW_BF16X3) && KC == 32 && NW == 4;
  static_assert(!DMA || (KC / 2 == 16 && MEM == 16 && NIT == 1), "one record per thread, 16 pairs x 16 members");
  __shared__ float4 recbuf[DMA ? NW * 5 * 64 : 1];
  auto block_sync = [&]() {
    if constexpr (DMA) lds_barrier();
    else __syncthreads();
  };
  const __amdgpu_buffer_rsrc_t rec_rs =  // (unused without DMA)
      __builtin_amdgcn_make_buffer_rsrc(const_cast<GT *>(Gb + (static_cast<size_t>(b0 + mb) * G.M1 + m) * G.Jh * 10), 0, 0x7fffffff, 0x00020000);
  auto dma_issue = [&](int c0) {
#pragma unroll
    for (int c = 0; c < 5; c++) {
      const int i = (wave * 5 + c) * 1024 + lane * 16;  // byte of the chunk: member i / 1280, pair (i % 1280) / 80
      const int j = i / 1280, o = i - j * 1280, r = o / 80;
      const int jl = min(j, nb - 1 - mb), jhc = min((c0 >> 1) + r, G.Jh - 1);  // (clamped pieces are never used)
      const unsigned off = (static_cast<unsigned>(jl) * G.M1 * G.Jh + jhc) * 80u + (o - r * 80);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rec_rs, (lds_void *)(recbuf + (wave * 5 + c) * 64), 16, off, 0, 0, 0);
    }
  };
  auto stage_load = [&](int c0) {
    if constexpr (PREC != LMW_FP64) {
      // registers: the raw records are kept until the chunk is written to LDS, so that nothing waits for them here
#pragma unroll
      for (int it = 0; it < NIT; it++) {
        rr_ok[it] = rec_ok(it, c0);
        const float4 *src = rec_src(it, c0);
#pragma unroll
        for (int c = 0; c < 5; c++) rr[it][c] = src[c];
      }
      return;
    }
    // FP64: the chunk's operands are MEM members x KC k x 4 complex values, each member's KC x 4 contiguous within a
    // part (64 bytes per latitude); request slot i of a thread takes piece P = (i NW + wave) 64 + lane, so that each
    // request of a wave reads 1 KB of one member's operands (one complex value per lane at a time touched 64 cache
    // lines per request).  Loads from clamped valid addresses; pok marks the pieces that are used.
#pragma unroll
    for (int i = 0; i < NP64; i++) {
      const int P = (i * NW + wave) * 64 + lane, j = P / (4 * KC), kl = (P >> 2) % KC, c = P & 3;
      const int k = c0 + kl, part = k >= G.Jhp ? 1 : 0, jh = part ? k - G.Jhp : k;
      pok[i] = P < MEM * KC * 4 && mb + j < nb && jh < G.Jh;
      const int jl = max(0, min(j, nb - 1 - mb)), jhc = min(jh, G.Jh - 1);
      pv[i] = Gb[(((static_cast<size_t>(b0 + mb + jl) * G.M1 + m) * 2 + p) * 2 + part) * 4 * G.Jh + static_cast<size_t>(jhc) * 4 + c];
    }
  };
  // split BF16 with 32-wide chunks: this chunk's table operands (FP32 entries for BF16X6, their two BF16 parts for
  // BF16X3) are requested as the chunk starts, ahead of the next chunk's records (loads retire in order, so the MFMAs
  // then wait for the tables only)
  constexpr bool PIPE = (PREC == LMW_BF16X6 || PREC == LMW_BF16X3) && KC == 32;
  v4f araw[TPW][2];
  // FP64: the table operands are requested FS64 k-steps at a time in the MFMA phase (the first FS64 of the first
  // tile before the staging, so that they have arrived when the MFMAs start)
  constexpr int FS64 = TPW > 1 ? 4 : 8;  // (two tiles per wave leave registers for four)
  double av0[FS64];
  if constexpr (DMA) dma_issue(0);
  else stage_load(0);
  for (int c0 = 0; c0 < KK; c0 += KC) {
    if constexpr (DMA) wait_vmem();  // this wave's pieces of the chunk's records (the only loads outstanding here)
    block_sync();                    // (and, for DMA, everyone's pieces)
    if constexpr (PREC == LMW_FP64) {
      const int r0 = rbase + 16 * min(wave, nrt - 1) + (lane & 15);  // (a wave without a tile loads a valid row)
#pragma unroll
      for (int i = 0; i < FS64; i++)
        av0[i] = T.A64[static_cast<size_t>(KK) * poff_m + static_cast<size_t>(c0 + 4 * i + (lane >> 4)) * kp + r0];
    }
    if constexpr (PIPE) {
#pragma unroll
      for (int ti = 0; ti < TPW; ti++) {
        const int t = min(wave + NW * ti, nrt - 1);  // (a wave without this tile loads a valid row it does not use)
        {
          const size_t ai = static_cast<size_t>(KK) * (poff_m + rbase + 16 * t + (lane & 15)) + c0 + 8 * (lane >> 4);
          if constexpr (PREC == LMW_BF16X6) {
            araw[ti][0] = *reinterpret_cast<const v4f *>(T.Abf + ai);
            araw[ti][1] = *reinterpret_cast<const v4f *>(T.Abf + ai + 4);
          } else {
            araw[ti][0] = *reinterpret_cast<const v4f *>(T.Abh + ai);
            araw[ti][1] = *reinterpret_cast<const v4f *>(T.Abl + ai);
          }
        }
      }
    }
    if constexpr (DMA) {
      // this chunk's records from the buffer (the table requests above stay in flight)
      const int jd = threadIdx.x >> 4, jp = threadIdx.x & 15;
      rr_ok[0] = mb + jd < nb && (c0 >> 1) + jp < G.Jh;
#pragma unroll
      for (int c = 0; c < 5; c++) rr[0][c] = recbuf[jd * 80 + jp * 5 + c];
    }
    if constexpr (PREC == LMW_FP64) {
      // operand (member j, k, value c) -> rows q = 2c (re), 2c + 1 (im) of the chunk, member column j ^ f64_swz(q, k):
      // a wave's 64 pieces then go to distinct LDS banks, and the MFMA's reads of a row stay one full row
#pragma unroll
      for (int i = 0; i < NP64; i++) {
        const int P = (i * NW + wave) * 64 + lane, j = P / (4 * KC), kl = (P >> 2) % KC, c = P & 3;
        if (P >= MEM * KC * 4) continue;
        const GT v = pok[i] ? pv[i] : GT{0, 0};
        const int col = j ^ f64_swz(2 * c, kl);
        bf[((2 * c) * KC + kl) * MEM + col] = v.x;
        bf[((2 * c + 1) * KC + kl) * MEM + col] = v.y;
      }
    }
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      if constexpr (PREC == LMW_FP64) break;
      const int e = threadIdx.x + 64 * NW * it;
      const int j = DMA ? e >> 4 : e % MEM, kk = DMA ? 2 * (e & 15) : 2 * (e / MEM);  // (member, first k of the pair)
      if (e >= (KC / 2) * MEM) continue;
      if constexpr (PREC != LMW_FP64) {
        // mixed: K runs over (latitude pair, P | H) interleaved, so the record of pair kk / 2 gives both its P
        // operands X (k = kk) and its H operands Y (k = kk + 1): parity 0 (even n - m) takes X from north + south
        // and Y from north - south, parity 1 the reverse (P(-mu) = (-1)^(n-m) P(mu), H(-mu) = -(-1)^(n-m) H(mu))
        const float sx = p ? -1.f : 1.f;
        const unsigned keep = rr_ok[it] ? 0xffffffffu : 0u;  // zero the lanes whose clamped record is not used
        auto kept = [&](float x) { return __builtin_bit_cast(float, __builtin_bit_cast(unsigned, x) & keep); };
        float4 r[5];
#pragma unroll
        for (int c = 0; c < 5; c++) r[c] = make_float4(kept(rr[it][c].x), kept(rr[it][c].y), kept(rr[it][c].z), kept(rr[it][c].w));
        const float dm = static_cast<float>(m);
        // north: A r0.xy, B r0.zw, UP r1.xy, VP r1.zw, E r2.xy; south: A r2.zw, B r3.xy, UP r3.zw, VP r4.xy, E r4.zw
        const float Ax = r[0].x + sx * r[2].z, Ay = r[0].y + sx * r[2].w;  // X combination (P rows)
        const float Bx = r[0].z + sx * r[3].x, By = r[0].w + sx * r[3].y;
        const float Ux = r[1].x + sx * r[3].z, Uy = r[1].y + sx * r[3].w;
        const float Ex = r[2].x + sx * r[4].z, Ey = r[2].y + sx * r[4].w;
        const float ax = r[0].x - sx * r[2].z, ay = r[0].y - sx * r[2].w;  // Y combination (H rows)
        const float bx = r[0].z - sx * r[3].x, by = r[0].w - sx * r[3].y;
        const float vx = r[1].z - sx * r[4].x, vy = r[1].w - sx * r[4].y;
        // X: -i m A, i m B, E, -i m U phi; Y: B, A, 0, V phi (sphere.cpp:374-383); i m z = (-m z.y, m z.x)
        const float X[8] = {dm * Ay, -(dm * Ax), -(dm * By), dm * Bx, Ex, Ey, dm * Uy, -(dm * Ux)};
        const float Y[8] = {bx, by, ax, ay, 0.f, 0.f, vx, vy};
#pragma unroll
        for (int q = 0; q < 8; q++) {
          vv[it][0][q] = X[q];
          vv[it][1][q] = Y[q];
        }
      }
#pragma unroll
      for (int q = 0; q < 8; q++) {
        if constexpr (PREC == LMW_FP32 || PREC == LMW_FP64) {
          bf[(q * KC + kk) * MEM + j] = vv[it][0][q];
          bf[(q * KC + kk + 1) * MEM + j] = vv[it][1][q];
        } else {
          unsigned p1, p2, p3;
          if constexpr (PREC == LMW_BF16X6) split3_pair(vv[it][0][q], vv[it][1][q], p1, p2, p3);
          else split2_pair(vv[it][0][q], vv[it][1][q], p1, p2);
          *reinterpret_cast<unsigned *>(bh + (q * MEM + j) * AP + kk) = p1;
          *reinterpret_cast<unsigned *>(blo + (q * MEM + j) * AP + kk) = p2;
          if constexpr (PREC == LMW_BF16X6) *reinterpret_cast<unsigned *>(b3p + (q * MEM + j) * AP + kk) = p3;
        }
      }
    }
    if constexpr (!DMA) stage_load(c0 + KC < KK ? c0 + KC : c0);  // (the last chunk requests itself again: the same load count)
    block_sync();
    // every wave has read its records: the next chunk's go to the buffer during this chunk's MFMAs
    if constexpr (DMA) dma_issue(c0 + KC < KK ? c0 + KC : c0);  // (the last chunk requests itself again: the same load count)
#pragma unroll
    for (int ti = 0; ti < TPW; ti++) {
      const int t = wave + NW * ti;
      if (t >= nrt) break;
      const int r = rbase + 16 * t + (lane & 15);
      const int kc = min(KC, KK - c0);  // KK is a multiple of 32
      if constexpr (PREC == LMW_FP64) {
        // the tile's table operands of this chunk are requested FS k-steps at a time, so the MFMAs wait for one
        // round trip per FS k-steps instead of one per k-step (kc == KC: KK is a multiple of 32)
        auto mma_steps = [&](const double (&av)[FS64], int s0) {
#pragma unroll
          for (int i = 0; i < FS64; i++) {
            const int kk = 4 * (s0 + i) + (lane >> 4);
            // o3 (q = 4, 5: the Laplacian part of ddelta) has no H term, so its operands in the H half of K are zeros
            // (Jhp is a multiple of 16: a k-step of 4 lies in one half)
            const bool hpart = c0 + 4 * (s0 + i) >= G.Jhp;
#pragma unroll
            for (int q = 0; q < 8; q++) {
              if ((q == 4 || q == 5) && hpart) continue;
              accv[ti][q] = mma64(av[i], bf[(q * KC + kk) * MEM + ((lane & 15) ^ f64_swz(q, kk))], accv[ti][q]);
            }
          }
        };
        int sb = 0;
        if (ti == 0) {  // the first k-steps of the first tile: requested before the staging
          mma_steps(av0, 0);
          sb = FS64;
        }
#pragma unroll 1
        for (int s0 = sb; s0 < KC / 4; s0 += FS64) {
          double av[FS64];
#pragma unroll
          for (int i = 0; i < FS64; i++)
            av[i] = T.A64[static_cast<size_t>(KK) * poff_m + static_cast<size_t>(c0 + 4 * (s0 + i) + (lane >> 4)) * kp + r];
          mma_steps(av, s0);
        }
        (void)kc;
      } else if constexpr (PREC == LMW_FP32) {
        // table operands requested 8 k-steps at a time (kc is a multiple of 32), one round trip per 8 k-steps
        for (int s0 = 0; s0 < kc / 4; s0 += 8) {
          float av[8];
#pragma unroll
          for (int i = 0; i < 8; i++)
            av[i] = T.A32[static_cast<size_t>(KK) * poff_m + static_cast<size_t>(c0 + 4 * (s0 + i) + (lane >> 4)) * kp + r];
#pragma unroll
          for (int i = 0; i < 8; i++) {
            const int kk = 4 * (s0 + i) + (lane >> 4);
#pragma unroll
            for (int q = 0; q < 8; q++) accv[ti][q] = mma32(av[i], bf[(q * KC + kk) * MEM + (lane & 15)], accv[ti][q]);
          }
        }
      } else {
        for (int s = 0; s < kc / 32; s++) {
          const size_t ai = static_cast<size_t>(KK) * (poff_m + r) + c0 + 32 * s + 8 * (lane >> 4);
          if constexpr (PREC == LMW_BF16X6) {
            v8bf ah, al, a3;
            if constexpr (PIPE) {
              split8_regs(araw[ti][0], araw[ti][1], ah, al, a3);
            } else {
              split8(T.Abf + ai, ah, al, a3);
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
              const int bo = (q * MEM + (lane & 15)) * AP + 32 * s + 8 * (lane >> 4);
              accv[ti][q] = mma6(ah, al, a3, load8(bh + bo), load8(blo + bo), load8(b3p + bo), accv[ti][q]);
            }
          } else {
            v8bf ah, al;
            if constexpr (PIPE) {
              ah = __builtin_bit_cast(v8bf, araw[ti][0]);
              al = __builtin_bit_cast(v8bf, araw[ti][1]);
            } else {
              ah = load8(T.Abh + ai);
              al = load8(T.Abl + ai);
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
              const int bo = (q * MEM + (lane & 15)) * AP + 32 * s + 8 * (lane >> 4);
              accv[ti][q] = mma3(ah, al, load8(bh + bo), load8(blo + bo), accv[ti][q]);
            }
          }
        }
      }
    }
  }
  if constexpr (DMA) wait_vmem();  // the last chunk's repeated request into the LDS buffer
  // Mixed-precision RK4 epilogue, two row tiles per wave: with NW < nrt row tiles the first nrt - NW waves own a
  // second tile and would update 8 elements per lane against 4 elsewhere, each element one memory round trip.  The
  // second tiles' accumulators go through the free operand chunk instead, and wave w updates its own 4 first-tile
  // elements plus list entries w x .. w x + x - 1 (x = nrt - NW) of the second tiles' elements g (tile NW + g / 4,
  // register g % 4): nrt elements for every wave.  (Truncations up to 170: nrt <= 6, so x <= 2 and the list takes
  // at most 16 KB of FP32 or 32 KB of FP64 accumulators, within every operand chunk.)  FP64 too: its 16x16x4 MFMA
  // interleaves the rows (register r of lane l holds row l / 16 + 4 r), which the list's k follows.
  constexpr bool BAL = TPW == 2;
  const int xlist = nrt - NW;
  ST *xl = reinterpret_cast<ST *>(lds_f);  // [second tile][r][q][lane]
  if (BAL && mode == ANAL_TEND && xlist > 0) {  // (block-uniform)
    block_sync();  // every wave is done with the operand chunk
    if (wave < xlist) {
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 8; q++) xl[((wave * 4 + r) * 8 + q) * 64 + lane] = accv[TPW - 1][q][r];
    }
    block_sync();
  }
  const int j = lane & 15, bl = mb + j;
  if (bl >= nb) return;
  const int b = b0 + bl, Km = G.N - m + 1;
  // this lane's elements e = 4 ti + r: row 16 t + 4 (lane / 16) + r of the parity block, t = wave + NW ti (the
  // FP64 MFMA interleaves: row 16 t + lane / 16 + 4 r)
  auto elem_k = [&](int e) {
    const int row = PREC == LMW_FP64 ? (lane >> 4) + 4 * (e & 3) : 4 * (lane >> 4) + (e & 3);
    return 2 * (16 * (wave + NW * (e >> 2)) + row) + p;
  };
  auto elem_ok = [&](int e) { return wave + NW * (e >> 2) < nrt && elem_k(e) < Km; };
  if (BAL && mode == ANAL_TEND) {
    // the first tile's 4 elements from registers, then this wave's list entries from the operand chunk, the state
    // loads of each element issued before the update of the one before; phi_00 P_0^0 of the stage input requested
    // with the first element's state
    auto list_k = [&](int g) {
      const int row = PREC == LMW_FP64 ? (lane >> 4) + 4 * (g & 3) : 4 * (lane >> 4) + (g & 3);
      return 2 * (16 * (NW + (g >> 2)) + row) + p;
    };
    const int nl = max(xlist, 0), g0 = wave * nl;
    const double phi00 = mixed ? xs[G.at(b, 2, 0)].x : 0.0;
    TendIn next = {};
    if (elem_ok(0)) next = tend_load(G, stage, b, moff_m + elem_k(0), x0, xs, acc, mixed);
    const double phibar = phi00 * 0.70710678118654752440;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const TendIn cur = next;
      if (e < 3) {
        if (elem_ok(e + 1)) next = tend_load(G, stage, b, moff_m + elem_k(e + 1), x0, xs, acc, mixed);
      } else if (nl > 0 && list_k(g0) < Km) {
        next = tend_load(G, stage, b, moff_m + list_k(g0), x0, xs, acc, mixed);
      }
      if (!elem_ok(e)) continue;
      const double2 o1 = cx(accv[0][0][e], accv[0][1][e]), o2 = cx(accv[0][2][e], accv[0][3][e]);
      const double2 o3 = cx(accv[0][4][e], accv[0][5][e]), o4 = cx(accv[0][6][e], accv[0][7][e]);
      tend_store(G, stage, dt, nu4, b, m + elem_k(e), moff_m + elem_k(e), o1, o2, o3, o4, cur, phibar, x0, xs, acc, mixed);
    }
    for (int i = 0; i < nl; i++) {
      const int g = g0 + i, k = list_k(g);
      const TendIn cur = next;
      if (i + 1 < nl && list_k(g + 1) < Km) next = tend_load(G, stage, b, moff_m + list_k(g + 1), x0, xs, acc, mixed);
      if (k >= Km) continue;
      const ST *xv = xl + g * 8 * 64 + lane;
      const double2 o1 = cx(xv[0], xv[64]), o2 = cx(xv[128], xv[192]), o3 = cx(xv[256], xv[320]), o4 = cx(xv[384], xv[448]);
      tend_store(G, stage, dt, nu4, b, m + k, moff_m + k, o1, o2, o3, o4, cur, phibar, x0, xs, acc, mixed);
    }
    return;
  }
  if (mode == ANAL_TEND) {
    // the RK4 update with the state loads of element e + 1 issued before the arithmetic and stores of element e
    // phi_00 P_0^0 of the stage input, requested with the first element's state
    const double phi00 = mixed ? xs[G.at(b, 2, 0)].x : 0.0;
    TendIn next = {};
    if (elem_ok(0)) next = tend_load(G, stage, b, moff_m + elem_k(0), x0, xs, acc, mixed);
    const double phibar = phi00 * 0.70710678118654752440;
#pragma unroll
    for (int e = 0; e < 4 * TPW; e++) {
      const TendIn cur = next;
      if (e + 1 < 4 * TPW && elem_ok(e + 1)) next = tend_load(G, stage, b, moff_m + elem_k(e + 1), x0, xs, acc, mixed);
      if (!elem_ok(e)) continue;
      const int ti = e >> 2, r = e & 3;
      const double2 o1 = cx(accv[ti][0][r], accv[ti][1][r]), o2 = cx(accv[ti][2][r], accv[ti][3][r]);
      const double2 o3 = cx(accv[ti][4][r], accv[ti][5][r]), o4 = cx(accv[ti][6][r], accv[ti][7][r]);
      tend_store(G, stage, dt, nu4, b, m + elem_k(e), moff_m + elem_k(e), o1, o2, o3, o4, cur, phibar, x0, xs, acc, mixed);
    }
    return;
  }
#pragma unroll
  for (int e = 0; e < 4 * TPW; e++) {
    if (!elem_ok(e)) continue;
    const int ti = e >> 2, r = e & 3;
    const double2 o1 = cx(accv[ti][0][r], accv[ti][1][r]), o2 = cx(accv[ti][2][r], accv[ti][3][r]);
    const double2 o3 = cx(accv[ti][4][r], accv[ti][5][r]), o4 = cx(accv[ti][6][r], accv[ti][7][r]);
    anal_epilogue(G, mode, stage, dt, nu4, b, bl, m + elem_k(e), moff_m + elem_k(e), o1, o2, o3, o4, x0, xs, acc, raw, mixed);
  }
}

// The FP64 inverse transform on the FP64 MFMA (16x16x4).  Block: one m, 8 members, one wave per 16 northern
// latitudes.  The MFMA's 16 columns are the (member, re/im) pairs of one complex quantity (the P and H tables are
// real), so a quantity takes one accumulator tile: P with zeta, phi, psi, chi and H with psi, chi, per parity
// (sphere.cpp:150-162,179-206), folded into the north and south rows at the end.  Rows of one parity are staged in
// LDS RC at a time as [quantity][row][2 member + re/im].  The FP64 MFMA holds row l/16 + 4r of its tile in lane l,
// register r; the two components of a complex result sit in neighbouring lanes.
template <int RC>
__global__ __launch_bounds__(512) void k_synth_f64(Geo G, const double *__restrict__ Ps, const double *__restrict__ Hs,
                                                   const double2 *__restrict__ spec, const double *__restrict__ knt,
                                                   double2 *__restrict__ Fs, int b0, int nb) {
  constexpr int MP = 8;
  __shared__ double cs[4][RC][16];
  const int m = blockIdx.x, mb = blockIdx.y * MP;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int Km = G.N - m + 1, kpe = G.kpe[m], kpo = G.kpo[m];
  const size_t moff_m = static_cast<size_t>(G.moff[m]);
  const int poff_m = G.poff[m];
  const double a = G.a;
  const int jt = wave;
  const bool active = jt * 16 < G.Jh;
  v4d accP[2][4], accH[2][2];
#pragma unroll
  for (int p = 0; p < 2; p++) {
#pragma unroll
    for (int q = 0; q < 4; q++) accP[p][q] = v4d{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 2; q++) accH[p][q] = v4d{0, 0, 0, 0};
  }
#pragma unroll
  for (int p = 0; p < 2; p++) {
    const int kp = p ? kpo : kpe;  // rows of this parity (16-padded)
    const int rbase = poff_m + (p ? kpe : 0);
    for (int c0 = 0; c0 < kp; c0 += RC) {
      __syncthreads();
      const int steps = min(RC, kp - c0) / 4;
      const int jh = jt * 16 + (lane & 15);
      // the chunk's table operands are requested 8 k-steps at a time, so the MFMAs wait for one round trip per
      // 8 k-steps instead of one per k-step (rows beyond the parity block load row 0 of it, unused); the first 8
      // before the coefficients are staged, so that they have arrived when the MFMAs start (93 -> 73 us per launch
      // at T170 x 32)
      double aP[8], aH[8];
      auto table_load = [&](int s0) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const int kk = 4 * (s0 + i) + (lane >> 4);
          const size_t ai = static_cast<size_t>(rbase + (s0 + i < steps ? c0 + kk : 0)) * G.Jhp + jh;
          aP[i] = Ps[ai];
          aH[i] = Hs[ai];
        }
      };
      if (active) table_load(0);
      for (int e = threadIdx.x; e < RC * MP; e += blockDim.x) {
        const int j = e % MP, kk = e / MP, bl = mb + j, kq = c0 + kk, k = 2 * kq + p, n = m + k;
        double2 z = cx(0, 0), d = cx(0, 0), f = cx(0, 0);
        if (bl < nb && k < Km && kq < kp) {
          const size_t q = moff_m + k;
          z = spec[G.at(b0 + bl, 0, q)];
          d = spec[G.at(b0 + bl, 1, q)];
          f = spec[G.at(b0 + bl, 2, q)];
        }
        const double kn = knt[min(n, G.N + 1)];  // psi, chi = kn (zeta, delta): sphere.cpp:185-190
        cs[0][kk][2 * j] = z.x;
        cs[0][kk][2 * j + 1] = z.y;
        cs[1][kk][2 * j] = f.x;
        cs[1][kk][2 * j + 1] = f.y;
        cs[2][kk][2 * j] = z.x * kn;
        cs[2][kk][2 * j + 1] = z.y * kn;
        cs[3][kk][2 * j] = d.x * kn;
        cs[3][kk][2 * j + 1] = d.y * kn;
      }
      __syncthreads();
      if (!active) continue;
      for (int s0 = 0; s0 < steps; s0 += 8) {
        if (s0 > 0) table_load(s0);
#pragma unroll
        for (int i = 0; i < 8; i++) {
          if (s0 + i >= steps) break;
          const int kk = 4 * (s0 + i) + (lane >> 4);
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const double bq = cs[q][kk][lane & 15];
            accP[p][q] = mma64(aP[i], bq, accP[p][q]);
            if (q >= 2) accH[p][q - 2] = mma64(aH[i], bq, accH[p][q - 2]);
          }
        }
      }
    }
  }
  if (!active) return;
  const int col = lane & 15, ri = col & 1, bl = mb + (col >> 1);
  const double dm = m, sgn = ri ? 1.0 : -1.0;  // (i m x) has the component sgn m x' of the partner component x'
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int jh = jt * 16 + (lane >> 4) + 4 * r;
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const double sg = s == 0 ? 1.0 : -1.0;
      auto P = [&](int q) { return accP[0][q][r] + sg * accP[1][q][r]; };
      auto H = [&](int q) { return sg * accH[0][q][r] + accH[1][q][r]; };
      const double psPo = __shfl_xor(P(2), 1), chPo = __shfl_xor(P(3), 1);  // every lane takes part
      const double u = sgn * dm * chPo - H(0), v = sgn * dm * psPo + H(1);  // sphere.cpp:200-201
      if (jh >= G.Jh || bl >= nb) continue;
      const int jj = s == 0 ? G.Jh + jh : G.Jh - 1 - jh;
      double *o = reinterpret_cast<double *>(Fs + ((static_cast<size_t>(b0 + bl) * G.J + jj) * G.M1 + m) * 4) + ri;
      o[0] = u / a;
      o[2] = v / a;
      o[4] = P(0);
      o[6] = P(1);
    }
  }
}

// XCD-aware block order of the per-group Legendre kernels: a 1-D grid of ceil(items / 8) * 8 * tiles blocks, the
// member tiles of one item 8 block ids apart.  Blocks go to the 8 XCDs round robin and in id order, so the tiles of
// an item run on one XCD at nearly the same time and the second reads the item's table rows from that XCD's L2.
// Returns false for the padding blocks of the last group of 8.
__device__ inline bool xcd_item(int items, int tiles, int &item, int &tile) {
  const int id = blockIdx.x, group = 8 * tiles;
  item = (id / group) * 8 + id % 8;
  tile = (id / 8) % tiles;
  return item < items;
}

// The inverse transform's order: chunks of 4 consecutive m on one XCD (chunk c on XCD c mod 8), the member tiles of
// an m 4 block ids apart within its XCD.  Its output records (32 bytes per (member, latitude, m)) of 4 consecutive m
// share a cache line, which then fills in one L2.  A 1-D grid of ceil(ceil(M1 / 4) / 8) * 8 * 4 * tiles blocks.
__device__ inline bool xcd_item4(int items, int tiles, int &item, int &tile) {
  const int x = blockIdx.x & 7, s = blockIdx.x >> 3, w = s % (4 * tiles);
  item = ((s / (4 * tiles)) * 8 + x) * 4 + (w & 3);
  tile = w >> 2;
  return item < items;
}

template <int PREC, int RC>
__global__ __launch_bounds__(512) void k_synth_mx(Geo G, SynthTabs T, const double2 *__restrict__ spec, float2 *__restrict__ Fs, int b0,
                                                  int nb, int mixed, int tiles) {
  int m, t;
  if (!xcd_item4(G.M1, tiles * kSynthJSplit, m, t)) return;
  // member tile t / kSynthJSplit; the block's waves take the row tiles of latitude part t % kSynthJSplit
  synth_mx_block<PREC, RC>(G, T, spec, Fs, b0, nb, mixed, m, (t / kSynthJSplit) * MEM, (t % kSynthJSplit) * (blockDim.x >> 6));
}

template <int PREC, int KC, int NW>
__global__ __launch_bounds__(512) void k_anal_mx(Geo G, AnalTabs T, const AnalIn<PREC> *__restrict__ Gb, int mode, int stage, double dt, double nu4,
                                                 double2 *__restrict__ x0, double2 *__restrict__ xs, double2 *__restrict__ acc,
                                                 double2 *__restrict__ raw, int b0, int nb, int mixed, int tiles) {
  int m, pt;
  if (!xcd_item(G.M1, 2 * tiles, m, pt)) return;  // the two parities and the member tiles of one m on one XCD
  anal_mx_block<PREC, KC, NW>(G, T, Gb, mode, stage, dt, nu4, x0, xs, acc, raw, b0, nb, mixed, m, pt & 1, (pt >> 1) * MEM);
}

// ------------------------------------------------------------------------------------------- multi-group kernels
// Persistent kernels over the truncation groups of an ensemble: a fixed grid of blocks walks a static work list
// (block w takes items w, w + gridDim.x, ...; the list is sorted by decreasing cost), so one launch per stage phase
// serves every active group and the small truncations fill the gaps of the large ones.  No counters: the schedule
// is fixed when the list is built.

struct SynthGroup {
  Geo G;
  SynthTabs T;
  const double2 *spec;
  float2 *Fs;
  int nb;
};
struct MultiSynth {
  SynthGroup g[kMaxGroups];
  const int4 *work;   // (group, m, member tile start, first latitude tile)
  const int *start;   // block x walks work[start[x] .. start[x + 1])
};
// blocks of kMultiWaves waves: an item covers kMultiWaves latitude tiles of one (m, member tile), as the per-group
// kernel's latitude parts do
constexpr int kMultiWaves = 4;
template <int PREC, int RC>
__global__ __launch_bounds__(64 * kMultiWaves, 3) void k_synth_mx_multi(MultiSynth M) {
  for (int w = M.start[blockIdx.x]; w < M.start[blockIdx.x + 1]; w++) {
    const int4 it = M.work[w];
    const SynthGroup &g = M.g[it.x];
    synth_mx_block<PREC, RC>(g.G, g.T, g.spec, g.Fs, 0, g.nb, 1, it.y, it.z, it.w);
  }
}

struct AnalGroup {
  Geo G;
  AnalTabs T;
  const float2 *Gb;
  double2 *x0, *xs, *acc;
  double dt, nu4;
  int nb;
};
struct MultiAnal {
  AnalGroup g[kMaxGroups];
  const int4 *work;  // (group, m, parity, member tile start)
  const int *start;  // block x walks work[start[x] .. start[x + 1])
  int stage;
};
template <int PREC, int KC, int NW>
__global__ __launch_bounds__(512, NW == 4 ? 3 : 1) void k_anal_mx_multi(MultiAnal M) {
  for (int w = M.start[blockIdx.x]; w < M.start[blockIdx.x + 1]; w++) {
    const int4 it = M.work[w];
    const AnalGroup &g = M.g[it.x];
    anal_mx_block<PREC, KC, NW, false>(g.G, g.T, g.Gb, ANAL_TEND, M.stage, g.dt, g.nu4, g.x0, g.xs, g.acc, nullptr, 0, g.nb, 1, it.y, it.z, it.w);
  }
}

// ------------------------------------------------------------------------------------------- host side

void SphereCore::build_mixed_tables() {
  if (mixed_ready) return;
  Geo G = geo();
  std::vector<int> mrow(R), soff(M1), ke32(M1), ko32(M1);
  int tot = 0;
  for (int m = 0; m <= N; m++) {
    for (int r = 0; r < h_kpe[m] + h_kpo[m]; r++) mrow[h_poff[m] + r] = m;
    const int Km = N - m + 1;
    ke32[m] = ((Km + 1) / 2 + 31) / 32 * 32;
    ko32[m] = (Km / 2 + 31) / 32 * 32;
    soff[m] = tot;
    tot += Jhp * (ke32[m] + ko32[m]);
  }
  d_mrow.alloc(R);
  d_mrow.upload(mrow.data(), R, 0, stream);
  d_soff.alloc(M1);
  d_soff.upload(soff.data(), M1, 0, stream);
  d_ke32.alloc(M1);
  d_ke32.upload(ke32.data(), M1, 0, stream);
  d_ko32.alloc(M1);
  d_ko32.upload(ko32.data(), M1, 0, stream);
  const size_t na = static_cast<size_t>(2 * Jhp) * R;
  A32.alloc(na);
  Abh.alloc(na);
  Abl.alloc(na);
  Ps32.alloc(static_cast<size_t>(R) * Jhp);
  Hs32.alloc(static_cast<size_t>(R) * Jhp);
  Fs32.alloc(Fs.n);
  Gb32.alloc(Gb.n);
  Sph.alloc(tot);
  Spl.alloc(tot);
  Shh.alloc(tot);
  Shl.alloc(tot);
  Spf.alloc(tot);
  Shf.alloc(tot);
  Abf.alloc(na);
  hipLaunchKernelGGL(k_build_anal_tables, dim3(1024), dim3(256), 0, stream, G, R, Pa.p, Ha.p, A32.p, Abh.p, Abl.p, Abf.p, d_mrow.p);
  LMW_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_build_synth_tables, dim3(M1), dim3(256), 0, stream, G, R, Ps.p, Hs.p, Ps32.p, Hs32.p, d_soff.p, d_ke32.p, d_ko32.p,
                     Sph.p, Spl.p, Shh.p, Shl.p, Spf.p, Shf.p);
  LMW_LAUNCH_CHECK();
  sync();
  mixed_ready = true;
}

// tile variants: synthesis row chunk RC; analysis (KC, NW)
const int kSynthRC[kSynthTiles] = {32, 64};
const int kAnalKC[kAnalTiles] = {32, 64, 128, 32, 64, 128};
const int kAnalNW[kAnalTiles] = {4, 4, 4, 8, 8, 8};

template <int PREC, int RC>
static void synth_mx_launch(SphereCore &c, const double2 *spec, int b0, int nb) {
  SynthTabs T{c.Ps32.p, c.Hs32.p, c.Sph.p, c.Spl.p, c.Shh.p, c.Shl.p, c.Spf.p, c.Shf.p, c.d_soff.p, c.d_ke32.p, c.d_ko32.p, c.knt.p};
  const int waves = ((c.Jh + 15) / 16 + kSynthJSplit - 1) / kSynthJSplit;  // row tiles of one latitude part
  const int tiles = (nb + MEM - 1) / MEM;
  dim3 grd(((c.M1 + 3) / 4 + 7) / 8 * 8 * 4 * tiles * kSynthJSplit), blk(64 * waves);
  const size_t lds = PREC == LMW_FP32 ? 8 * RC * MEM * sizeof(float) : nsplit(PREC) * 8 * (RC + 8) * MEM * sizeof(__bf16);
  hipLaunchKernelGGL((k_synth_mx<PREC, RC>), grd, blk, lds, c.stream, c.geo(), T, spec, c.Fs32.p, b0, nb, 1, tiles);
  LMW_LAUNCH_CHECK();
}

void SphereCore::launch_synth_mx(const double2 *spec, int b0, int nb) {
  const int rc = kSynthRC[tile_synth];
  if (precision == LMW_FP32) {
    if (rc == 64) synth_mx_launch<LMW_FP32, 64>(*this, spec, b0, nb);
    else synth_mx_launch<LMW_FP32, 32>(*this, spec, b0, nb);
  } else if (precision == LMW_BF16X6) {
    if (rc == 64) synth_mx_launch<LMW_BF16X6, 64>(*this, spec, b0, nb);
    else synth_mx_launch<LMW_BF16X6, 32>(*this, spec, b0, nb);
  } else {
    if (rc == 64) synth_mx_launch<LMW_BF16X3, 64>(*this, spec, b0, nb);
    else synth_mx_launch<LMW_BF16X3, 32>(*this, spec, b0, nb);
  }
}

template <int PREC, int KC, int NW>
static void anal_mx_launch(SphereCore &c, int mode, int stage, double dt, int b0, int nb, double2 *raw_out) {
  AnalTabs T{c.A32.p, c.Abh.p, c.Abl.p, c.Abf.p, c.A64.p};
  const int tiles = (nb + MEM - 1) / MEM;
  dim3 grd((c.M1 + 7) / 8 * 8 * 2 * tiles), blk(64 * NW);
  const size_t lds = PREC == LMW_FP64   ? 8 * KC * MEM * sizeof(double)
                     : PREC == LMW_FP32 ? 8 * KC * MEM * sizeof(float)
                                        : nsplit(PREC) * 8 * (KC + 8) * MEM * sizeof(__bf16);
  const AnalIn<PREC> *gb;
  if constexpr (PREC == LMW_FP64) gb = c.Gb.p;
  else gb = c.Gb32.p;
  hipLaunchKernelGGL((k_anal_mx<PREC, KC, NW>), grd, blk, lds, c.stream, c.geo(), T, gb, mode, stage, dt, c.prm.diffusion, c.x0.p, c.xs.p,
                     c.acc.p, raw_out, b0, nb, PREC == LMW_FP64 ? 0 : 1, tiles);
  LMW_LAUNCH_CHECK();
}

// The FP64 inverse transform on the FP64 MFMA (FP64 synthesis variants kFp64Tiles and kFp64Tiles + 1: RC = 32, 64)
void SphereCore::launch_synth_fp64_mfma(int rc, const double2 *spec, int b0, int nb, bool scratch) {
  Geo G = geo();
  if (scratch) {  // member-major scratch buffers
    G.sb = 3LL * S;
    G.sf = S;
    G.sq = 1;
  }
  const int waves = (Jh + 15) / 16;
  dim3 grd(M1, (nb + 7) / 8), blk(64 * waves);
  if (rc == 64) hipLaunchKernelGGL(k_synth_f64<64>, grd, blk, 0, stream, G, Ps.p, Hs.p, spec, knt.p, Fs.p, b0, nb);
  else hipLaunchKernelGGL(k_synth_f64<32>, grd, blk, 0, stream, G, Ps.p, Hs.p, spec, knt.p, Fs.p, b0, nb);
  LMW_LAUNCH_CHECK();
}

// The FP64 step's forward transform on the FP64 MFMA (FP64 analysis variants kFp64Tiles and kFp64Tiles + 1: KC = 32
// and 4 or 8 waves).  Its table is built when such a variant is chosen, never inside a captured graph.
void SphereCore::build_fp64_mfma_tables() {
  if (A64.n) return;
  A64.alloc(static_cast<size_t>(2 * Jhp) * R);
  hipLaunchKernelGGL(k_build_anal64, dim3(M1), dim3(256), 0, stream, geo(), Pa.p, Ha.p, A64.p);
  LMW_LAUNCH_CHECK();
  sync();
}
void SphereCore::launch_anal_fp64_mfma(int nw, int mode, int stage, double dt, int b0, int nb, double2 *raw_out) {
  if (!A64.n) runtime("the FP64 MFMA tables were not built before the launch");
  if (nw == 8) anal_mx_launch<LMW_FP64, 32, 8>(*this, mode, stage, dt, b0, nb, raw_out);
  else anal_mx_launch<LMW_FP64, 32, 4>(*this, mode, stage, dt, b0, nb, raw_out);
}

template <int PREC>
static void anal_mx_dispatch(SphereCore &c, int mode, int stage, double dt, int b0, int nb, double2 *raw_out) {
  switch (c.tile_anal) {
    case 0: anal_mx_launch<PREC, 32, 4>(c, mode, stage, dt, b0, nb, raw_out); break;
    case 2: anal_mx_launch<PREC, 128, 4>(c, mode, stage, dt, b0, nb, raw_out); break;
    case 3: anal_mx_launch<PREC, 32, 8>(c, mode, stage, dt, b0, nb, raw_out); break;
    case 4: anal_mx_launch<PREC, 64, 8>(c, mode, stage, dt, b0, nb, raw_out); break;
    case 5: anal_mx_launch<PREC, 128, 8>(c, mode, stage, dt, b0, nb, raw_out); break;
    default: anal_mx_launch<PREC, 64, 4>(c, mode, stage, dt, b0, nb, raw_out); break;
  }
}

void SphereCore::launch_anal_mx(int mode, int stage, double dt, int b0, int nb, double2 *raw_out) {
  if (precision == LMW_FP32) anal_mx_dispatch<LMW_FP32>(*this, mode, stage, dt, b0, nb, raw_out);
  else if (precision == LMW_BF16X6) anal_mx_dispatch<LMW_BF16X6>(*this, mode, stage, dt, b0, nb, raw_out);
  else anal_mx_dispatch<LMW_BF16X3>(*this, mode, stage, dt, b0, nb, raw_out);
}

// ------------------------------------------------------------------------------------------- multi-group host side

// resident blocks of `kernel` on the whole device
template <class K> static int resident_blocks(K kernel, int threads, size_t lds) {
  int per_cu = 0, dev = 0, cus = 0;
  LMW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
  LMW_HIP(hipGetDevice(&dev));
  LMW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return std::max(1, per_cu) * cus;
}

static size_t synth_multi_lds(int prec, int rc) {
  return prec == LMW_FP32 ? 8 * rc * MEM * sizeof(float) : nsplit(prec) * 8 * (rc + 8) * MEM * sizeof(__bf16);
}
static size_t anal_multi_lds(int prec, int kc) {
  return prec == LMW_FP32 ? 8 * kc * MEM * sizeof(float) : nsplit(prec) * 8 * (kc + 8) * MEM * sizeof(__bf16);
}
template <int PREC> static int synth_multi_blocks(int rc) {
  const size_t lds = synth_multi_lds(PREC, rc);
  return rc == 64 ? resident_blocks(k_synth_mx_multi<PREC, 64>, 64 * kMultiWaves, lds)
                  : resident_blocks(k_synth_mx_multi<PREC, 32>, 64 * kMultiWaves, lds);
}
template <int PREC> static int anal_multi_blocks(int tile) {
  const int kc = kAnalKC[tile], nw = kAnalNW[tile];
  const size_t lds = anal_multi_lds(PREC, kc);
  switch (tile) {
    case 0: return resident_blocks(k_anal_mx_multi<PREC, 32, 4>, 64 * nw, lds);
    case 2: return resident_blocks(k_anal_mx_multi<PREC, 128, 4>, 64 * nw, lds);
    case 3: return resident_blocks(k_anal_mx_multi<PREC, 32, 8>, 64 * nw, lds);
    case 4: return resident_blocks(k_anal_mx_multi<PREC, 64, 8>, 64 * nw, lds);
    case 5: return resident_blocks(k_anal_mx_multi<PREC, 128, 8>, 64 * nw, lds);
    default: return resident_blocks(k_anal_mx_multi<PREC, 64, 4>, 64 * nw, lds);
  }
}

// Static schedule of one launch: the items of one key (same group and m, or 4 consecutive m for the inverse transform,
// whose output records of 4 consecutive m share cache lines) go to one XCD -- block x runs on XCD x mod 8, the
// hardware dealing blocks to the XCDs round robin -- so they share that XCD's L2 for the table rows and operands; keys
// go to the least loaded XCD in decreasing order of cost, then each XCD's items, longest first, to its least loaded
// block.  Costs estimate an item's duration from its chunk count (the K loop) and its epilogue.
struct SchedItem {
  int4 v;
  int key;
  double cost;
};
static void schedule(std::vector<SchedItem> items, int nblk, DevArray<int4> &work, DevArray<int> &start) {
  std::map<int, double> key_cost;
  for (const auto &it : items) key_cost[it.key] += it.cost;
  std::vector<std::pair<double, int>> keys;
  for (const auto &kc : key_cost) keys.push_back({kc.second, kc.first});
  std::stable_sort(keys.begin(), keys.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
  const int nx = std::min(8, nblk);
  std::vector<double> xload(nx, 0.0);
  std::map<int, int> key_xcd;
  for (const auto &k : keys) {
    const int x = static_cast<int>(std::min_element(xload.begin(), xload.end()) - xload.begin());
    xload[x] += k.first;
    key_xcd[k.second] = x;
  }
  std::stable_sort(items.begin(), items.end(), [](const SchedItem &a, const SchedItem &b) { return a.cost > b.cost; });
  std::vector<std::vector<int4>> lists(nblk);
  std::vector<double> bload(nblk, 0.0);
  for (const auto &it : items) {
    const int x = key_xcd[it.key];
    int best = -1;
    for (int blk = x; blk < nblk; blk += nx)
      if (best < 0 || bload[blk] < bload[best]) best = blk;
    bload[best] += it.cost;
    lists[best].push_back(it.v);
  }
  std::vector<int4> flat;
  std::vector<int> off(nblk + 1, 0);
  for (int blk = 0; blk < nblk; blk++) {
    off[blk] = static_cast<int>(flat.size());
    flat.insert(flat.end(), lists[blk].begin(), lists[blk].end());
  }
  off[nblk] = static_cast<int>(flat.size());
  work.alloc(std::max<size_t>(flat.size(), 1));
  start.alloc(nblk + 1);
  if (!flat.empty()) LMW_HIP(hipMemcpy(work.p, flat.data(), flat.size() * sizeof(int4), hipMemcpyHostToDevice));
  LMW_HIP(hipMemcpy(start.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
}

void build_multi_work(const std::vector<SphereCore *> &groups, unsigned active_mask, MultiWork &w) {
  const SphereCore &big = *groups.back();
  const int rc = kSynthRC[big.tile_synth], kc = kAnalKC[big.tile_anal], nw = kAnalNW[big.tile_anal];
  switch (big.precision) {
    case LMW_FP32: w.synth_blocks = synth_multi_blocks<LMW_FP32>(rc); w.anal_blocks = anal_multi_blocks<LMW_FP32>(big.tile_anal); break;
    case LMW_BF16X6: w.synth_blocks = synth_multi_blocks<LMW_BF16X6>(rc); w.anal_blocks = anal_multi_blocks<LMW_BF16X6>(big.tile_anal); break;
    default: w.synth_blocks = synth_multi_blocks<LMW_BF16X3>(rc); w.anal_blocks = anal_multi_blocks<LMW_BF16X3>(big.tile_anal); break;
  }
  std::vector<SchedItem> sy, an;
  for (int g = 0; g < static_cast<int>(groups.size()); g++) {
    if (!(active_mask & (1u << g))) continue;
    const SphereCore &c = *groups[g];
    const int lat_tiles = (c.Jh + 15) / 16, parts = (lat_tiles + kMultiWaves - 1) / kMultiWaves;
    for (int m = 0; m <= c.N; m++) {
      const int Km = c.N - m + 1;
      const int ke32 = ((Km + 1) / 2 + 31) / 32 * 32, ko32 = (Km / 2 + 31) / 32 * 32;
      const int chunks = (ke32 + rc - 1) / rc + (ko32 + rc - 1) / rc;
      for (int mb = 0; mb < c.B; mb += MEM) {
        for (int pt = 0; pt < parts; pt++)
          sy.push_back({make_int4(g, m, mb, pt * kMultiWaves), g * 4096 + m / 4, 4.6 * chunks * (rc / 32) + 3.0});
        for (int p = 0; p < 2; p++) {
          const int rows = p ? c.h_kpo[m] : c.h_kpe[m];
          if (rows == 0) continue;
          const int tpw = (rows / 16 + nw - 1) / nw;  // row tiles of the busiest wave
          an.push_back({make_int4(g, m, p, mb), g * 4096 + m, 2.9 * (2 * c.Jhp / kc) * (kc / 32) + 6.4 * tpw + 2.0});
        }
      }
    }
  }
  w.nsynth = static_cast<int>(sy.size());
  w.nanal = static_cast<int>(an.size());
  schedule(sy, w.synth_blocks, w.synth, w.synth_start);
  schedule(an, w.anal_blocks, w.anal, w.anal_start);
}

template <int PREC, int RC>
static void synth_multi(const std::vector<SphereCore *> &gs, const MultiWork &w, hipStream_t stream) {
  MultiSynth M{};
  for (size_t g = 0; g < gs.size(); g++) {
    SphereCore &c = *gs[g];
    M.g[g] = SynthGroup{c.geo(), SynthTabs{c.Ps32.p, c.Hs32.p, c.Sph.p, c.Spl.p, c.Shh.p, c.Shl.p, c.Spf.p, c.Shf.p, c.d_soff.p, c.d_ke32.p,
                                             c.d_ko32.p, c.knt.p},
                        c.xs.p, c.Fs32.p, c.B};
  }
  M.work = w.synth.p;
  M.start = w.synth_start.p;
  hipLaunchKernelGGL((k_synth_mx_multi<PREC, RC>), dim3(w.synth_blocks), dim3(64 * kMultiWaves), synth_multi_lds(PREC, RC), stream, M);
  LMW_LAUNCH_CHECK();
}

void launch_synth_mx_multi(const std::vector<SphereCore *> &gs, const MultiWork &w, hipStream_t stream) {
  if (w.nsynth == 0) return;
  const SphereCore &big = *gs.back();
  const bool rc64 = kSynthRC[big.tile_synth] == 64;
  switch (big.precision) {
    case LMW_FP32: rc64 ? synth_multi<LMW_FP32, 64>(gs, w, stream) : synth_multi<LMW_FP32, 32>(gs, w, stream); break;
    case LMW_BF16X6: rc64 ? synth_multi<LMW_BF16X6, 64>(gs, w, stream) : synth_multi<LMW_BF16X6, 32>(gs, w, stream); break;
    default: rc64 ? synth_multi<LMW_BF16X3, 64>(gs, w, stream) : synth_multi<LMW_BF16X3, 32>(gs, w, stream); break;
  }
}

template <int PREC, int KC, int NW>
static void anal_multi(const std::vector<SphereCore *> &gs, const std::vector<double> &dts, int stage, const MultiWork &w, hipStream_t stream) {
  MultiAnal M{};
  for (size_t g = 0; g < gs.size(); g++) {
    SphereCore &c = *gs[g];
    M.g[g] = AnalGroup{c.geo(), AnalTabs{c.A32.p, c.Abh.p, c.Abl.p, c.Abf.p, nullptr}, c.Gb32.p, c.x0.p, c.xs.p, c.acc.p, dts[g], c.prm.diffusion,
                       c.B};
  }
  M.work = w.anal.p;
  M.start = w.anal_start.p;
  M.stage = stage;
  hipLaunchKernelGGL((k_anal_mx_multi<PREC, KC, NW>), dim3(w.anal_blocks), dim3(64 * NW), anal_multi_lds(PREC, KC), stream, M);
  LMW_LAUNCH_CHECK();
}

template <int PREC>
static void anal_multi_dispatch(const std::vector<SphereCore *> &gs, const std::vector<double> &dts, int stage, const MultiWork &w,
                                hipStream_t stream) {
  switch (gs.back()->tile_anal) {
    case 0: anal_multi<PREC, 32, 4>(gs, dts, stage, w, stream); break;
    case 2: anal_multi<PREC, 128, 4>(gs, dts, stage, w, stream); break;
    case 3: anal_multi<PREC, 32, 8>(gs, dts, stage, w, stream); break;
    case 4: anal_multi<PREC, 64, 8>(gs, dts, stage, w, stream); break;
    case 5: anal_multi<PREC, 128, 8>(gs, dts, stage, w, stream); break;
    default: anal_multi<PREC, 64, 4>(gs, dts, stage, w, stream); break;
  }
}

void launch_anal_mx_multi(const std::vector<SphereCore *> &gs, const std::vector<double> &dts, int stage, const MultiWork &w,
                          hipStream_t stream) {
  if (w.nanal == 0) return;
  switch (gs.back()->precision) {
    case LMW_FP32: anal_multi_dispatch<LMW_FP32>(gs, dts, stage, w, stream); break;
    case LMW_BF16X6: anal_multi_dispatch<LMW_BF16X6>(gs, dts, stage, w, stream); break;
    default: anal_multi_dispatch<LMW_BF16X3>(gs, dts, stage, w, stream); break;
  }
}

}  // namespace lmw
