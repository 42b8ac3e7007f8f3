// Epilogue of the forward Legendre transform, shared by the FP64 and the mixed-precision kernels: the outputs
// o1..o4 of one (member, m, n) become the state (set_state_grid), raw transforms, or the tendencies and the RK4
// stage update of cpp/sphere.cpp:387-389 and :407-436 with the del^4 filter of :392-405.
#pragma once
#include "sphere_core.h"

namespace lmw {

// The RK4 stage update (mode ANAL_TEND) in two halves, so that a kernel can issue the loads of its next element
// before the arithmetic and stores of the current one: tend_load reads the stage's inputs of one (member b, packed
// index q), tend_store applies sphere.cpp:387-389, :392-405 and :407-436 to them.  phibar = phi_00 / sqrt 2 of the
// stage input (mixed precision only).
struct TendIn {
  double2 s0[3], a[3], d;
};
__device__ inline TendIn tend_load(const Geo &G, int stage, int b, size_t q, const double2 *__restrict__ x0, const double2 *__restrict__ xs,
                                   const double2 *__restrict__ acc, int mixed) {
  TendIn t;
#pragma unroll
  for (int fld = 0; fld < 3; fld++) {
    const size_t i = G.at(b, fld, q);
    t.s0[fld] = x0[i];
    t.a[fld] = stage >= 2 ? acc[i] : cx(0, 0);
  }
  t.d = mixed ? xs[G.at(b, 1, q)] : cx(0, 0);
  return t;
}
// (n, q): degree and packed index of the coefficient; callers compute q from a moff[m] read before any store, since the
// compiler cannot tell that the state stores leave the Geo tables unchanged and would read them again after each store
__device__ inline void tend_store(const Geo &G, int stage, double dt, double nu4, int b, int n, size_t q, double2 o1, double2 o2, double2 o3,
                                  double2 o4, const TendIn &t, double phibar, double2 *__restrict__ x0, double2 *__restrict__ xs,
                                  double2 *__restrict__ acc, int mixed) {
  double2 kz = o1, kd = cadd(o2, cscale(o3, n * (n + 1.0) / (G.a * G.a))), kf = o4;
  if (mixed) kf = csub(kf, cscale(t.d, phibar));
  if (q == 0) kz = kd = kf = cx(0, 0);  // sphere.cpp:387-389
  {
#pragma clang fp contract(off)
    const double2 kk[3] = {kz, kd, kf};
    double fv = 1.0, fpp = 1.0;
    if (stage == 4 && nu4 > 0 && n >= 1) {  // del^4 filter (sphere.cpp:392-405)
      const double a2 = G.a * G.a, l = n * (n + 1.0);
      const double kv = (l - 2) / a2, kq = l / a2;
      fv = exp(-nu4 * kv * kv * dt);
      fpp = exp(-nu4 * kq * kq * dt);
    }
    // The (0, 0) coefficients have zero tendency and no filter, so every stage leaves their state unchanged: only
    // their RK4 sum is written.  Nothing in the launch rewrites xs at q = 0, so every block may read phi_00 of the
    // stage input (the mean height of the mixed path's phibar) without a race.
    const bool keep = q == 0;
#pragma unroll
    for (int fld = 0; fld < 3; fld++) {
      const size_t i = G.at(b, fld, q);
      const double2 s0 = t.s0[fld], kv = kk[fld];
      if (keep) {
        if (stage <= 3) acc[i] = stage == 1 ? kv : cadd(t.a[fld], cscale(kv, 2.0));
        continue;
      }
      if (stage == 1) {
        acc[i] = kv;
        xs[i] = cadd(s0, cscale(kv, 0.5 * dt));
      } else if (stage == 2) {
        acc[i] = cadd(t.a[fld], cscale(kv, 2.0));
        xs[i] = cadd(s0, cscale(kv, 0.5 * dt));
      } else if (stage == 3) {
        acc[i] = cadd(t.a[fld], cscale(kv, 2.0));
        xs[i] = cadd(s0, cscale(kv, dt));
      } else {
        double2 nx = cadd(s0, cscale(cadd(t.a[fld], kv), dt / 6));
        if (stage == 4 && nu4 > 0 && n >= 1) nx = cscale(nx, fld == 2 ? fpp : fv);
        x0[i] = nx;
        xs[i] = nx;
      }
    }
  }
}

// mixed != 0: the phi-flux operands were formed with phi - phibar (phibar = phi_00 / sqrt 2, the area mean), and the
// mean part -phibar D(U, V) = -phibar delta is added here from the stage input delta, exactly.
__device__ inline void anal_epilogue(const Geo &G, int mode, int stage, double dt, double nu4, int b, int bl, int n, size_t q, double2 o1,
                                     double2 o2, double2 o3, double2 o4, double2 *__restrict__ x0, double2 *__restrict__ xs,
                                     double2 *__restrict__ acc, double2 *__restrict__ raw, int mixed) {
  const size_t S = G.S;
  if (mode == ANAL_RAW) {  // raw outputs are indexed by the local member
    raw[(static_cast<size_t>(bl) * 4 + 0) * S + q] = o1;
    raw[(static_cast<size_t>(bl) * 4 + 1) * S + q] = o2;
    raw[(static_cast<size_t>(bl) * 4 + 2) * S + q] = o3;
    raw[(static_cast<size_t>(bl) * 4 + 3) * S + q] = o4;
    return;
  }
  if (mode == ANAL_STATE) {
    if (q == 0) o1 = o2 = cx(0, 0);  // the constant is not defined by the winds (sphere.cpp:213-214)
    x0[G.at(b, 0, q)] = o1;
    x0[G.at(b, 1, q)] = o2;
    x0[G.at(b, 2, q)] = o4;
    xs[G.at(b, 0, q)] = o1;
    xs[G.at(b, 1, q)] = o2;
    xs[G.at(b, 2, q)] = o4;
    return;
  }
  if (mode == ANAL_TEND) {
    const double phibar = mixed ? xs[G.at(b, 2, 0)].x * 0.70710678118654752440 : 0.0;  // phi_00 P_0^0
    tend_store(G, stage, dt, nu4, b, n, q, o1, o2, o3, o4, tend_load(G, stage, b, q, x0, xs, acc, mixed), phibar, x0, xs, acc, mixed);
    return;
  }
  // ANAL_TENDOUT: the tendencies only, indexed by the local member
  double2 kz = o1, kd = cadd(o2, cscale(o3, n * (n + 1.0) / (G.a * G.a))), kf = o4;
  if (mixed) {
    const double phibar = xs[G.at(b, 2, 0)].x * 0.70710678118654752440;  // phi_00 P_0^0
    kf = csub(kf, cscale(xs[G.at(b, 1, q)], phibar));
  }
  if (q == 0) kz = kd = kf = cx(0, 0);  // sphere.cpp:387-389
  raw[(static_cast<size_t>(bl) * 4 + 0) * S + q] = kz;
  raw[(static_cast<size_t>(bl) * 4 + 1) * S + q] = kd;
  raw[(static_cast<size_t>(bl) * 4 + 2) * S + q] = kf;
}

}  // namespace lmw
