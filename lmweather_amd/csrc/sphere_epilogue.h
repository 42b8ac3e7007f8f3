// Epilogue of the forward Legendre transform, shared by the FP64 and the mixed-precision kernels: the outputs
// o1..o4 of one (member, m, n) become the state (set_state_grid), raw transforms, or the tendencies and the RK4
// stage update of cpp/sphere.cpp:387-389 and :407-436 with the del^4 filter of :392-405.
#pragma once
#include "sphere_core.h"

namespace lmw {

// mixed != 0: the phi-flux operands were formed with phi - phibar (phibar = phi_00 / sqrt 2, the area mean), and the
// mean part -phibar D(U, V) = -phibar delta is added here from the stage input delta, exactly.
__device__ inline void anal_epilogue(const Geo &G, int mode, int stage, double dt, double nu4, int b, int bl, int m, int k, double2 o1, double2 o2,
                                     double2 o3, double2 o4, double2 *__restrict__ x0, double2 *__restrict__ xs, double2 *__restrict__ acc,
                                     double2 *__restrict__ raw, int mixed) {
  const int n = m + k;
  const size_t q = static_cast<size_t>(G.moff[m]) + k;
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
  double2 kz = o1, kd = cadd(o2, cscale(o3, n * (n + 1.0) / (G.a * G.a))), kf = o4;
  if (mixed) {
    const double phibar = xs[G.at(b, 2, 0)].x * 0.70710678118654752440;  // phi_00 P_0^0
    kf = csub(kf, cscale(xs[G.at(b, 1, q)], phibar));
  }
  if (q == 0) kz = kd = kf = cx(0, 0);  // sphere.cpp:387-389
  if (mode == ANAL_TENDOUT) {           // tendencies only, indexed by the local member
    raw[(static_cast<size_t>(bl) * 4 + 0) * S + q] = kz;
    raw[(static_cast<size_t>(bl) * 4 + 1) * S + q] = kd;
    raw[(static_cast<size_t>(bl) * 4 + 2) * S + q] = kf;
    return;
  }
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
    for (int fld = 0; fld < 3; fld++) {
      const size_t i = G.at(b, fld, q);
      const double2 s0 = x0[i], kv = kk[fld];
      if (stage == 1) {
        acc[i] = kv;
        xs[i] = cadd(s0, cscale(kv, 0.5 * dt));
      } else if (stage == 2) {
        acc[i] = cadd(acc[i], cscale(kv, 2.0));
        xs[i] = cadd(s0, cscale(kv, 0.5 * dt));
      } else if (stage == 3) {
        acc[i] = cadd(acc[i], cscale(kv, 2.0));
        xs[i] = cadd(s0, cscale(kv, dt));
      } else {
        double2 nx = cadd(s0, cscale(cadd(acc[i], kv), dt / 6));
        if (stage == 4 && nu4 > 0 && n >= 1) nx = cscale(nx, fld == 2 ? fpp : fv);
        x0[i] = nx;
        xs[i] = nx;
      }
    }
  }
}

}  // namespace lmw
