// Arakawa C-grid shallow water on an f/beta-plane channel for gfx950 (reference cpp/plane.cpp): the C ABI of the
// plane (include/lmw.h) and the FP64 instantiation of its device code (plane_core.h), with FMA contraction off.
#define LMW_PLANE_CONTRACT _Pragma("clang fp contract(off)")
#include "plane_core.h"

namespace lmw {
std::unique_ptr<PlaneCoreBase> make_plane_core_f64(const lmw_plane_params &p, int members) {
  return std::make_unique<PlaneCore<double>>(p, members);
}
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
    if (batch > 65535) invalid("a plane batch holds at most 65535 members");
    auto h = std::make_unique<lmw_plane>();
    if (precision == LMW_FP64) h->core = make_plane_core_f64(p, batch);
    else if (precision == LMW_FP32) h->core = make_plane_core_f32(p, batch);
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
    pl.core->perturb(member, amplitude, seed);
  });
}
lmw_status lmw_plane_fields(lmw_plane *p, int member, const char *h, const char *u, const char *v) {
  return guard([&] {
    lmw_plane &pl = P(p);
    check_member(pl, member);
    NEEDP(h);
    NEEDP(u);
    NEEDP(v);
    pl.core->fields_expr(member, h, u, v);
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
lmw_status lmw_plane_steps_timed(lmw_plane *p, double dt, int64_t count, double *kernel_ms) {
  return guard([&] {
    lmw_plane &pl = P(p);
    NEEDP(kernel_ms);
    if (count < 0) invalid("step count must not be negative");
    pl.core->steps(dt, count, kernel_ms);
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
lmw_status lmw_plane_get_params(lmw_plane *p, lmw_plane_params *params, int *batch, int *precision) {
  return guard([&] {
    lmw_plane &pl = P(p);
    if (params) *params = pl.core->prm;
    if (batch) *batch = pl.core->B;
    if (precision) *precision = pl.precision;
  });
}

// Zonal harmonics (plane.cpp:264-310).  c_m of an h row is taken at the cell centres (shift 1/2), of a u row at
// the west faces (shift 0), of a v row at the cell centres in x (shift 1/2).
namespace {
void check_harmonic(const lmw_plane &pl, int m, int j, int jmin, const char *what) {
  if (m < 1 || m >= pl.core->prm.nx / 2 || j < jmin || j >= pl.core->prm.ny) invalid(what);
}
// the x in [0, lx/m) of a crest of 2|c| cos(2 pi m x / lx + arg c) (plane.cpp:276-284)
double crest_x(const double *c, int m, double lx) {
  constexpr double kPi = 3.14159265358979323846;
  const double ph = std::atan2(c[1], c[0]);
  double x = -ph / (2 * kPi * m) * lx;
  const double per = lx / m;
  x = std::fmod(x, per);
  return x < 0 ? x + per : x;
}
}  // namespace
lmw_status lmw_plane_zonal_amplitude(lmw_plane *p, int member, int m, int j, double *out) {
  return guard([&] {
    lmw_plane &pl = P(p);
    check_member(pl, member);
    NEEDP(out);
    check_harmonic(pl, m, j, 0, "zonal harmonic or row out of range");
    double c[2];
    pl.core->harmonic(member, LMW_PLANE_H, m, 1, j, 1, c);
    *out = 2 * std::hypot(c[0], c[1]);
  });
}
lmw_status lmw_plane_zonal_crest(lmw_plane *p, int member, int m, int j, double *out) {
  return guard([&] {
    lmw_plane &pl = P(p);
    check_member(pl, member);
    NEEDP(out);
    check_harmonic(pl, m, j, 0, "zonal harmonic or row out of range");
    double c[2];
    pl.core->harmonic(member, LMW_PLANE_H, m, 1, j, 1, c);
    *out = crest_x(c, m, pl.core->prm.lx);
  });
}
lmw_status lmw_plane_v_zonal_crest(lmw_plane *p, int member, int m, int j, double *out) {
  return guard([&] {
    lmw_plane &pl = P(p);
    check_member(pl, member);
    NEEDP(out);
    check_harmonic(pl, m, j, 1, "zonal harmonic or v row out of range");
    double c[2];
    pl.core->harmonic(member, LMW_PLANE_V, m, 1, j, 1, c);
    *out = crest_x(c, m, pl.core->prm.lx);
  });
}
// sum over the u rows and the interior v rows of |c_m|^2, over ny (plane.cpp:286-299)
lmw_status lmw_plane_harmonic_energy(lmw_plane *p, int member, int m, double *out) {
  return guard([&] {
    lmw_plane &pl = P(p);
    check_member(pl, member);
    NEEDP(out);
    const int ny = pl.core->prm.ny;
    if (m < 1 || m >= pl.core->prm.nx / 2) invalid("zonal harmonic out of range");
    std::vector<double> cu(2 * static_cast<size_t>(ny)), cv(2 * static_cast<size_t>(ny - 1));
    pl.core->harmonic(member, LMW_PLANE_U, m, 0, 0, ny, cu.data());
    pl.core->harmonic(member, LMW_PLANE_V, m, 1, 1, ny - 1, cv.data());
    double s = 0;
    for (int j = 0; j < ny; j++) s += cu[2 * j] * cu[2 * j] + cu[2 * j + 1] * cu[2 * j + 1];
    for (int j = 0; j < ny - 1; j++) s += cv[2 * j] * cv[2 * j] + cv[2 * j + 1] * cv[2 * j + 1];
    *out = s / ny;
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
