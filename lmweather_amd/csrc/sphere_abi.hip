// C ABI of the sphere and of ensembles (include/lmw.h): the reference's Sphere methods (cpp/sphere.cpp) on top of
// SphereCore, with the reference's argument checks, error texts, run() splitting and periodic checks.  Host code
// here only converts layouts and evaluates O(N^2) point diagnostics; every O(J K N) transform runs on the GPU.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <memory>

#include "sphere_core.h"

namespace lmw {

namespace {
thread_local std::string g_err;
std::atomic<int64_t> g_bytes{0};

double eps_host(int n, int m) { return std::sqrt(static_cast<double>(n * n - m * m) / static_cast<double>(4 * n * n - 1)); }

// legendre_column (sphere.cpp:66-73) on the host, for the O(N^2) point evaluators
void legendre_column_host(int m, int nmax, double mu, std::vector<double> &out) {
  out.assign(static_cast<size_t>(nmax + 2 - m), 0.0);
  double c = std::sqrt(std::max(0.0, 1 - mu * mu)), p = 1 / std::sqrt(2.0);
  for (int k = 1; k <= m; k++) p *= std::sqrt((2.0 * k + 1) / (2.0 * k)) * c;
  out[0] = p;
  if (nmax + 1 >= m + 1) out[1] = std::sqrt(2.0 * m + 3) * mu * p;
  for (int n = m + 2; n <= nmax + 1; n++) out[n - m] = (mu * out[n - 1 - m] - eps_host(n - 1, m) * out[n - 2 - m]) / eps_host(n, m);
}

uint64_t splitmix(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
}  // namespace

void set_error(const std::string &msg) { g_err = msg; }
void note_bytes(int64_t delta) { g_bytes += delta; }

// One member of a SphereCore seen through the reference's Sphere interface.
struct MemberView {
  SphereCore *c;
  int b;
  int N() const { return c->N; }
  size_t ngrid() const { return static_cast<size_t>(c->J) * c->K; }
  size_t nfull() const { return static_cast<size_t>(c->N + 1) * (c->N + 1); }

  // reference layout (N+1)^2 <-> packed triangle
  void to_packed(const double *re_im, std::vector<double2> &packed) const {
    packed.assign(c->S, make_double2(0, 0));
    for (int m = 0; m <= N(); m++)
      for (int n = m; n <= N(); n++) {
        const size_t i = static_cast<size_t>(m) * (N() + 1) + n;
        packed[c->packed_index(m, n)] = make_double2(re_im[2 * i], re_im[2 * i + 1]);
      }
  }
  void from_packed(const std::vector<double2> &packed, double *re_im) const {
    std::fill(re_im, re_im + 2 * nfull(), 0.0);
    for (int m = 0; m <= N(); m++)
      for (int n = m; n <= N(); n++) {
        const size_t i = static_cast<size_t>(m) * (N() + 1) + n;
        const double2 v = packed[c->packed_index(m, n)];
        re_im[2 * i] = v.x;
        re_im[2 * i + 1] = v.y;
      }
  }
  std::vector<double2> spectral(int field) const {
    std::vector<double2> p(c->S);
    if (field == LMW_SPEC_Q) {  // the tracer lives in its own buffer (single-member cores)
      if (!c->has_tracer) runtime("no tracer: set one with tracer() first");
      c->q0.download(p.data(), c->S, 0, c->stream);
      c->sync();
      return p;
    }
    c->get_spectral_packed(b, field, p.data());
    return p;
  }
  void set_spectral(int field, const std::vector<double2> &p) { c->set_spectral_packed(b, field, p.data()); }

  // stable_dt (sphere.cpp:438-447)
  double stable_dt(double safety) {
    c->compute_grids(b, 1);
    c->reduce(b, 1);
    double r[16];
    c->download_red(b, r);
    const double wmax = r[3], hmax = r[4];
    if (!(hmax > 0)) runtime("the height field must be positive to step the model");
    // advect-only: only the tracer moves, at the frozen wind speed (sphere.cpp:443-444)
    const double omega = c->advect_only ? (c->N / c->prm.radius) * wmax
                                        : (c->N / c->prm.radius) * (std::sqrt(c->prm.gravity * hmax) + wmax) + 2 * c->prm.omega;
    if (!(omega > 0)) return 1e9;
    return safety / omega;
  }
  // check_finite (sphere.cpp:449-455)
  void check_finite() {
    double sum = c->spectral_abs_sum(b);
    if (c->has_tracer) sum += c->tracer_abs_sum();
    if (!std::isfinite(sum)) runtime("the integration blew up (not finite): use a smaller dt");
  }
  void diag_sums(double out[16]) {
    c->compute_grids(b, 1);
    c->reduce(b, 1);
    c->download_red(b, out);
  }
  double diag(int which) {
    double r[16];
    switch (which) {
      case LMW_DIAG_MEAN_HEIGHT: {  // sphere.cpp:510
        std::vector<double2> phi = spectral(2);
        return phi[0].x / std::sqrt(2.0) / c->prm.gravity;
      }
      case LMW_DIAG_ENERGY: diag_sums(r); return r[0] + r[1];
      case LMW_DIAG_KINETIC: diag_sums(r); return r[0];
      case LMW_DIAG_POTENTIAL: diag_sums(r); return r[1];
      case LMW_DIAG_ENSTROPHY: diag_sums(r); return r[2];
      case LMW_DIAG_MAX_WIND: diag_sums(r); return r[3];
      case LMW_DIAG_ERROR_L2: diag_sums(r); return std::sqrt(r[5] / r[6]);
      case LMW_DIAG_ERROR_LINF: diag_sums(r); return r[7] / r[8];
      case LMW_DIAG_TIME: return c->t[b];
      case LMW_DIAG_TRACER_MEAN: {  // sphere.cpp:271-274
        std::vector<double2> q = spectral(3);
        return q[0].x / std::sqrt(2.0);
      }
      case LMW_DIAG_TRACER_MASS: return tracer_mass();
      case LMW_DIAG_TRACER_MIN:
      case LMW_DIAG_TRACER_MAX: {  // sphere.cpp:284-285
        std::vector<double> g(ngrid());
        c->tracer_grid(g.data());
        return which == LMW_DIAG_TRACER_MIN ? *std::min_element(g.begin(), g.end()) : *std::max_element(g.begin(), g.end());
      }
      case LMW_DIAG_TRACER_ERROR_L2:
      case LMW_DIAG_TRACER_ERROR_LINF: {
        double l2, linf;
        tracer_errors(l2, linf);
        return which == LMW_DIAG_TRACER_ERROR_L2 ? l2 : linf;
      }
      default: invalid("unknown diagnostic");
    }
  }
  // tracer_mass (sphere.cpp:275-283): area mean of h q, row sums in the reference's order
  double tracer_mass() {
    if (!c->has_tracer) runtime("no tracer: set one with tracer() first");
    std::vector<double> h(ngrid()), q(ngrid()), w(c->J);
    grid(LMW_GRID_H, h.data());
    c->tracer_grid(q.data());
    c->w.download(w.data(), c->J, 0, c->stream);
    c->sync();
    double s = 0;
    for (int j = 0; j < c->J; j++) {
      double row = 0;
      for (int k = 0; k < c->K; k++) row += h[static_cast<size_t>(j) * c->K + k] * q[static_cast<size_t>(j) * c->K + k];
      s += w[j] * row / c->K;
    }
    return 0.5 * s;
  }
  // tracer_error_l2/linf (sphere.cpp:571-589 with g = 1 against the reference tracer grid)
  void tracer_errors(double &l2, double &linf) {
    if (!c->has_tracer) runtime("no tracer: set one with tracer() first");
    std::vector<double> q(ngrid()), r(ngrid()), w(c->J);
    c->tracer_grid(q.data());
    c->ref_q.download(r.data(), ngrid(), 0, c->stream);
    c->w.download(w.data(), c->J, 0, c->stream);
    c->sync();
    double num = 0, den = 0, mn = 0, mr = 0;
    for (int j = 0; j < c->J; j++)
      for (int k = 0; k < c->K; k++) {
        const size_t i = static_cast<size_t>(j) * c->K + k;
        const double rr = r[i] / 1.0, d = q[i] - rr;
        num += w[j] * d * d;
        den += w[j] * rr * rr;
        mn = std::max(mn, std::fabs(d));
        mr = std::max(mr, std::fabs(rr));
      }
    l2 = std::sqrt(num / den);
    linf = mn / mr;
  }
  void grid(int field, double *out) {
    if (field == LMW_GRID_TRACER) {
      c->tracer_grid(out);
      return;
    }
    if (field == LMW_GRID_OROGRAPHY) {  // the stored hs grid (sphere.hpp:50)
      const size_t plane = static_cast<size_t>(c->J) * c->K;
      c->hs.download(out, plane, static_cast<size_t>(c->hs_slot[b]) * plane, c->stream);
      c->sync();
      return;
    }
    if (field < 0 || field > 3) invalid("unknown grid field");
    c->compute_grids(b, 1);
    c->download_grid(b, field, out);
  }
  // spectrum (sphere.cpp:536-551)
  void spectrum(double *ke, double *pe) {
    std::vector<double2> z = spectral(0), d = spectral(1), f = spectral(2);
    const double a2 = c->prm.radius * c->prm.radius, g = c->prm.gravity, g2 = g * g;
    std::fill(ke, ke + N() + 1, 0.0);
    std::fill(pe, pe + N() + 1, 0.0);
    for (int m = 0; m <= N(); m++) {
      const double cm = m == 0 ? 1.0 : 2.0;
      for (int n = std::max(m, 1); n <= N(); n++) {
        const size_t q = c->packed_index(m, n);
        const double nz = z[q].x * z[q].x + z[q].y * z[q].y, nd = d[q].x * d[q].x + d[q].y * d[q].y;
        const double np = f[q].x * f[q].x + f[q].y * f[q].y;
        ke[n] += 0.25 * cm * a2 * (nz + nd) / (n * (n + 1.0));
        pe[n] += 0.25 * cm * g * np / g2;
      }
    }
  }
  // height_at / wind_at / vorticity_at (sphere.cpp:592-637)
  double point(int which, double lat, double lon) {
    const double mu = std::sin(lat), a = c->prm.radius;
    if (which == LMW_POINT_TRACER && !c->has_tracer) runtime("no tracer: set one with tracer() first");
    std::vector<double2> z = spectral(0), d = spectral(1), f = which == LMW_POINT_TRACER ? spectral(3) : spectral(2);
    std::vector<double> col;
    double sum = 0;
    for (int m = 0; m <= N(); m++) {
      legendre_column_host(m, N(), mu, col);
      double gr, gi;
      if (which == LMW_POINT_HEIGHT || which == LMW_POINT_VORTICITY || which == LMW_POINT_TRACER) {
        const std::vector<double2> &s = which == LMW_POINT_VORTICITY ? z : f;
        double sr = 0, si = 0;
        for (int n = m; n <= N(); n++) {
          const double2 v = s[c->packed_index(m, n)];
          sr += v.x * col[n - m];
          si += v.y * col[n - m];
        }
        gr = sr;
        gi = si;
      } else {
        double spr = 0, spi = 0, scr = 0, sci = 0, hpr = 0, hpi = 0, hcr = 0, hci = 0;
        for (int n = std::max(m, 1); n <= N(); n++) {
          const double k = -a * a / (n * (n + 1.0));
          const double P = col[n - m];
          const double H = -n * eps_host(n + 1, m) * col[n + 1 - m] + (n > m ? (n + 1) * eps_host(n, m) * col[n - 1 - m] : 0.0);
          const double2 ps = make_double2(z[c->packed_index(m, n)].x * k, z[c->packed_index(m, n)].y * k);
          const double2 ch = make_double2(d[c->packed_index(m, n)].x * k, d[c->packed_index(m, n)].y * k);
          spr += ps.x * P; spi += ps.y * P; scr += ch.x * P; sci += ch.y * P;
          hpr += ps.x * H; hpi += ps.y * H; hcr += ch.x * H; hci += ch.y * H;
        }
        if (which == LMW_POINT_U) { gr = (-m * sci - hpr) / a; gi = (m * scr - hpi) / a; }
        else { gr = (-m * spi + hcr) / a; gi = (m * spr + hci) / a; }
      }
      sum += (m == 0 ? 1.0 : 2.0) * (gr * std::cos(m * lon) - gi * std::sin(m * lon));
    }
    if (which == LMW_POINT_HEIGHT) return sum / c->prm.gravity;
    if (which == LMW_POINT_U || which == LMW_POINT_V) return sum / std::cos(lat);
    return sum;
  }
  // run (sphere.cpp:457-484)
  void run(double seconds, double dt) {
    if (!(seconds >= 0)) invalid("run: seconds must not be negative");
    if (c->advect_only && !c->has_tracer) invalid("advect_only: set a tracer first (Sphere.tracer)");
    if (seconds == 0) return;
    if (!(dt > 0)) dt = stable_dt(1.5);
    const double limit = stable_dt(2.0 * std::sqrt(2.0));
    if (!(dt <= limit))
      invalid("run: dt = " + fstr(dt) + " s exceeds the RK4 stability limit of about " + fstr(limit) + " s for the present state: use a smaller dt");
    int64_t n = static_cast<int64_t>(std::ceil(seconds / dt - 1e-9));
    if (n < 1) n = 1;
    if (n > 50000000) invalid("run: more than 5e7 steps; use a larger dt");
    const double h = seconds / static_cast<double>(n);
    for (int64_t i = 0; i < n;) {
      const int64_t chunk = std::min<int64_t>(64 - (i & 63), n - i);  // up to the next (i & 63) == 63 step
      c->steps_graph(h, chunk);
      for (int64_t q = 0; q < chunk; q++) c->t[b] += h;
      i += chunk;
      if (((i - 1) & 63) == 63) {
        check_finite();
        const double now = stable_dt(2.0 * std::sqrt(2.0));
        if (!(h <= now))
          runtime("run: dt = " + fstr(h) + " s exceeds the RK4 stability limit of about " + fstr(now) + " s reached at t = " + fstr(c->t[b]) +
                  " s (the flow accelerated): use a smaller dt");
      }
    }
    check_finite();
  }
  // perturb (sphere.cpp:313-336): the seeded stream and its scaling on the host, the synthesis on the GPU
  void perturb(double amplitude, uint64_t seed, int max_wave) {
    if (max_wave < 1 || max_wave > N()) invalid("max_wave must be between 1 and the truncation");
    uint64_t st = seed * 0x9E3779B97F4A7C15ULL + 12345;
    auto normal = [&] {
      double s = 0;
      for (int i = 0; i < 12; i++) s += static_cast<double>(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
      return s - 6.0;
    };
    std::vector<double2> add(c->S, make_double2(0, 0));
    for (int m = 0; m <= max_wave; m++)
      for (int n = std::max(m, 1); n <= max_wave; n++) {
        const double re = normal(), im = m == 0 ? 0.0 : normal();
        add[c->packed_index(m, n)] = make_double2(re, im);
      }
    std::vector<double> g(ngrid()), mu(c->J), w(c->J);
    c->synthesize_raw(nullptr, nullptr, add.data(), nullptr, nullptr, nullptr, g.data());
    c->mu.download(mu.data(), c->J, 0, c->stream);
    c->w.download(w.data(), c->J, 0, c->stream);
    c->sync();
    double ms = 0, wsum = 0;
    for (int j = 0; j < c->J; j++)
      for (int k = 0; k < c->K; k++) {
        const double x = g[static_cast<size_t>(j) * c->K + k];
        ms += w[j] * x * x;
        wsum += w[j];
      }
    const double rms = std::sqrt(ms / wsum);
    const double scale = rms > 0 ? amplitude / rms * c->prm.gravity : 0;
    std::vector<double2> phi = spectral(2);
    for (int i = 0; i < c->S; i++) phi[i] = make_double2(phi[i].x + add[i].x * scale, phi[i].y + add[i].y * scale);
    set_spectral(2, phi);
    c->t[b] = 0;
    c->mark_reference(b);
    c->sync();
  }
  // add_mode (sphere.cpp:301-311)
  void add_mode(int n, int m, double amplitude, double mean_depth) {
    if (m < 0 || n < std::max(m, 1) || n > N()) invalid("mode (n, m) needs 0 <= m <= n <= truncation and n >= 1");
    std::vector<double2> zero(c->S, make_double2(0, 0)), phi(c->S, make_double2(0, 0));
    phi[0] = make_double2(c->prm.gravity * mean_depth * std::sqrt(2.0), 0);
    const size_t q = c->packed_index(m, n);
    phi[q].x += c->prm.gravity * amplitude * (m == 0 ? 1.0 : 0.5);
    set_spectral(0, zero);
    set_spectral(1, zero);
    set_spectral(2, phi);
    c->t[b] = 0;
    c->mark_reference(b);
    c->sync();
  }
};

}  // namespace lmw

using namespace lmw;

struct lmw_sphere {
  std::unique_ptr<SphereCore> core;
  MemberView view() { return MemberView{core.get(), 0}; }
};

namespace lmw {
std::vector<SphereCore *> all_groups(lmw_ensemble *e);
void drop_macro(lmw_ensemble *e);
}  // namespace lmw

struct lmw_ensemble {
  std::vector<std::unique_ptr<SphereCore>> groups;  // ascending truncation
  std::vector<std::pair<int, int>> where;           // member -> (group, local index)
  // macro step: every group advanced by its own sub-steps, captured as one HIP graph whose branches are the groups'
  // streams (forked from and joined to `origin`)
  hipStream_t origin = nullptr;
  hipGraphExec_t macro = nullptr;       // one macro step
  hipGraphExec_t macro_many = nullptr;  // kMacroGraph macro steps (branch mode: each group's steps in one chain)
  static constexpr int kMacroGraph = 4;
  std::vector<double> macro_dts;
  std::vector<int> macro_sub;
  std::vector<hipEvent_t> events;  // fork, then one join per group
  // persistent mode: the mixed Legendre transforms of all active groups in one multi-group launch per stage phase
  bool persistent = false;
  std::map<unsigned, std::unique_ptr<MultiWork>> work;  // per set of active groups (bit mask)
  ~lmw_ensemble() {
    if (macro) (void)hipGraphExecDestroy(macro);
    if (macro_many) (void)hipGraphExecDestroy(macro_many);
    for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
    if (origin) (void)hipStreamDestroy(origin);
  }
  MemberView view(int member) {
    if (member < 0 || member >= static_cast<int>(where.size())) invalid("member index out of range");
    return MemberView{groups[where[member].first].get(), where[member].second};
  }
};

#define NEED(p) \
  if (!(p)) invalid("null pointer argument")

extern "C" {

const char *lmw_last_error(void) { return g_err.c_str(); }

lmw_status lmw_device_info(char *name, size_t name_len, int *compute_units, size_t *bytes_in_use) {
  return guard([&] {
    hipDeviceProp_t prop;
    int dev = 0;
    LMW_HIP(hipGetDevice(&dev));
    LMW_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && name_len) {
      std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (bytes_in_use) *bytes_in_use = static_cast<size_t>(g_bytes.load());
  });
}

lmw_status lmw_fft(double *re_im, size_t n, int count, int sign) {
  return guard([&] {
    if (n < 2 || (n & (n - 1))) invalid("fft length must be a power of two");
    if (n > 2048) invalid("fft length must be at most 2048");
    if (count < 1) return;
    NEED(re_im);
    fft_batch(reinterpret_cast<double2 *>(re_im), static_cast<int>(n), count, sign);
  });
}

lmw_status lmw_gauss_legendre(int n, double *x, double *w) {
  return guard([&] {
    if (n < 1) invalid("n must be positive");
    NEED(x && w);
    gauss_device(n, x, w);
  });
}

// ------------------------------------------------------------------------------------------------ sphere
lmw_status lmw_sphere_create(int truncation, const lmw_sphere_params *params, lmw_sphere **out) {
  return guard([&] {
    NEED(out);
    *out = nullptr;
    lmw_sphere_params p = params ? *params : lmw_sphere_params{6.37122e6, 7.292e-5, 9.80616, 0.0};
    auto s = std::make_unique<lmw_sphere>();
    s->core = std::make_unique<SphereCore>(truncation, p, 1, LMW_FP64);
    *out = s.release();
  });
}
lmw_status lmw_sphere_destroy(lmw_sphere *s) {
  return guard([&] { delete s; });
}
lmw_status lmw_sphere_dims(const lmw_sphere *s, int *truncation, int *nlat, int *nlon) {
  return guard([&] {
    NEED(s);
    if (truncation) *truncation = s->core->N;
    if (nlat) *nlat = s->core->J;
    if (nlon) *nlon = s->core->K;
  });
}
lmw_status lmw_sphere_set_precision(lmw_sphere *s, int precision) {
  return guard([&] {
    NEED(s);
    s->core->set_precision(precision);
  });
}
lmw_status lmw_sphere_set_state_grid(lmw_sphere *s, const double *h, const double *u, const double *v, size_t n) {
  return guard([&] {
    NEED(s);
    if (n != s->view().ngrid()) invalid("grid size mismatch");
    NEED(h && u && v);
    s->core->set_state_grid(0, h, u, v);
    s->core->sync();
  });
}
lmw_status lmw_sphere_set_orography_grid(lmw_sphere *s, const double *hs, size_t n) {
  return guard([&] {
    NEED(s);
    if (n != s->view().ngrid()) invalid("grid size mismatch");
    NEED(hs);
    s->core->set_orography(0, hs);
    s->core->sync();
  });
}
lmw_status lmw_sphere_set_axis_tilt(lmw_sphere *s, double alpha) {
  return guard([&] {
    NEED(s);
    s->core->set_axis_tilt(0, alpha);
    s->core->sync();
  });
}
lmw_status lmw_sphere_set_diffusion(lmw_sphere *s, double nu4) {
  return guard([&] {
    NEED(s);
    s->core->prm.diffusion = nu4;
    s->core->drop_graph();
  });
}
lmw_status lmw_sphere_mark_reference(lmw_sphere *s) {
  return guard([&] {
    NEED(s);
    s->core->mark_reference(0);
    s->core->sync();
  });
}
lmw_status lmw_sphere_add_mode(lmw_sphere *s, int n, int m, double amplitude, double mean_depth) {
  return guard([&] {
    NEED(s);
    s->view().add_mode(n, m, amplitude, mean_depth);
  });
}
lmw_status lmw_sphere_setup(lmw_sphere *s, int which, double alpha) {
  return guard([&] {
    NEED(s);
    s->core->setup(0, 1, which, alpha);
  });
}
lmw_status lmw_sphere_fields(lmw_sphere *s, const char *h, const char *u, const char *v) {
  return guard([&] {
    NEED(s);
    NEED(h);
    NEED(u);
    NEED(v);
    s->core->fields_expr(0, 1, h, u, v);
  });
}
lmw_status lmw_sphere_orography(lmw_sphere *s, const char *hs) {
  return guard([&] {
    NEED(s);
    NEED(hs);
    s->core->orography_expr(0, 1, hs);
  });
}
lmw_status lmw_sphere_tracer(lmw_sphere *s, const char *q) {
  return guard([&] {
    NEED(s);
    NEED(q);
    s->core->tracer_expr(q);
  });
}
lmw_status lmw_sphere_perturb(lmw_sphere *s, double amplitude, uint64_t seed, int max_wave) {
  return guard([&] {
    NEED(s);
    s->view().perturb(amplitude, seed, max_wave);
  });
}
lmw_status lmw_sphere_step(lmw_sphere *s, double dt) {
  return guard([&] {
    NEED(s);
    s->core->steps(dt, 1);
    s->core->t[0] += dt;
    s->core->sync();
  });
}
lmw_status lmw_sphere_steps(lmw_sphere *s, double dt, int64_t count) {
  return guard([&] {
    NEED(s);
    if (count < 0) invalid("step count must not be negative");
    s->core->steps_graph(dt, count);
    for (int64_t i = 0; i < count; i++) s->core->t[0] += dt;
    s->core->sync();
  });
}
lmw_status lmw_sphere_run(lmw_sphere *s, double seconds, double dt) {
  return guard([&] {
    NEED(s);
    s->view().run(seconds, dt);
  });
}
lmw_status lmw_sphere_stable_dt(lmw_sphere *s, double safety, double *out) {
  return guard([&] {
    NEED(s && out);
    *out = s->view().stable_dt(safety);
  });
}
lmw_status lmw_sphere_get_grid(lmw_sphere *s, int field, double *out, size_t n) {
  return guard([&] {
    NEED(s);
    if (n != s->view().ngrid()) invalid("grid size mismatch");
    NEED(out);
    s->view().grid(field, out);
  });
}
lmw_status lmw_sphere_get_spectral(lmw_sphere *s, int field, double *re_im, size_t n_complex) {
  return guard([&] {
    NEED(s);
    MemberView v = s->view();
    if (field < 0 || field > 3) invalid("unknown spectral field");
    if (field == LMW_SPEC_Q && !s->core->has_tracer) runtime("no tracer: set one with tracer() first");
    if (n_complex != v.nfull()) invalid("spectral size mismatch");
    NEED(re_im);
    v.from_packed(v.spectral(field), re_im);
  });
}
lmw_status lmw_sphere_set_spectral(lmw_sphere *s, int field, const double *re_im, size_t n_complex) {
  return guard([&] {
    NEED(s);
    MemberView v = s->view();
    if (field < 0 || field > 3) invalid("unknown spectral field");
    if (n_complex != v.nfull()) invalid("spectral size mismatch");
    NEED(re_im);
    std::vector<double2> p;
    v.to_packed(re_im, p);
    if (field == LMW_SPEC_Q) s->core->set_tracer_spectral(p.data());  // a tracer from its coefficients
    else v.set_spectral(field, p);
    s->core->sync();
  });
}
lmw_status lmw_sphere_set_tracer_grid(lmw_sphere *s, const double *q, size_t n) {
  return guard([&] {
    NEED(s);
    if (n != s->view().ngrid()) invalid("grid size mismatch");
    NEED(q);
    s->core->set_tracer_grid(q);
  });
}
lmw_status lmw_sphere_set_advect_only(lmw_sphere *s, int on) {
  return guard([&] {
    NEED(s);
    s->core->set_advect_only(on != 0);
  });
}
lmw_status lmw_sphere_tracer_flags(const lmw_sphere *s, int *has_tracer, int *advect_only) {
  return guard([&] {
    NEED(s);
    if (has_tracer) *has_tracer = s->core->has_tracer ? 1 : 0;
    if (advect_only) *advect_only = s->core->advect_only ? 1 : 0;
  });
}
lmw_status lmw_sphere_diag(lmw_sphere *s, int which, double *out) {
  return guard([&] {
    NEED(s && out);
    *out = s->view().diag(which);
  });
}
lmw_status lmw_sphere_spectrum(lmw_sphere *s, double *ke, double *pe, size_t n) {
  return guard([&] {
    NEED(s);
    if (n != static_cast<size_t>(s->core->N + 1)) invalid("spectrum size mismatch");
    NEED(ke && pe);
    s->view().spectrum(ke, pe);
  });
}
lmw_status lmw_sphere_point(lmw_sphere *s, int which, double lat, double lon, double *out) {
  return guard([&] {
    NEED(s && out);
    if (which < 0 || which > LMW_POINT_TRACER) invalid("unknown point evaluator");
    *out = s->view().point(which, lat, lon);
  });
}
lmw_status lmw_sphere_mode_coefficient(lmw_sphere *s, int n, int m, int imag, double *out) {
  return guard([&] {
    NEED(s && out);
    if (m < 0 || n < m || n > s->core->N) invalid("mode (n, m) outside the truncation");
    const double2 c = s->view().spectral(2)[s->core->packed_index(m, n)];
    *out = (imag ? c.y : c.x) / s->core->prm.gravity;
  });
}
lmw_status lmw_sphere_wave_crest(lmw_sphere *s, int m, double lat, double *out) {
  return guard([&] {
    NEED(s && out);
    if (m < 1 || m > s->core->N) invalid("zonal wavenumber must be between 1 and the truncation");
    std::vector<double> col;
    legendre_column_host(m, s->core->N, std::sin(lat), col);
    std::vector<double2> phi = s->view().spectral(2);
    double gr = 0, gi = 0;
    for (int n = m; n <= s->core->N; n++) {
      gr += phi[s->core->packed_index(m, n)].x * col[n - m];
      gi += phi[s->core->packed_index(m, n)].y * col[n - m];
    }
    *out = -std::atan2(gi, gr) / m;
  });
}
lmw_status lmw_sphere_synthesize(lmw_sphere *s, const double *spec, double *grid) {
  return guard([&] {
    NEED(s && spec && grid);
    std::vector<double2> p;
    s->view().to_packed(spec, p);
    s->core->synthesize_raw(nullptr, nullptr, p.data(), nullptr, nullptr, nullptr, grid);
  });
}
lmw_status lmw_sphere_winds(lmw_sphere *s, const double *zeta, const double *delta, double *U, double *V) {
  return guard([&] {
    NEED(s && zeta && delta && U && V);
    std::vector<double2> pz, pd;
    s->view().to_packed(zeta, pz);
    s->view().to_packed(delta, pd);
    s->core->synthesize_raw(pz.data(), pd.data(), nullptr, U, V, nullptr, nullptr);
  });
}
lmw_status lmw_sphere_analyze(lmw_sphere *s, const double *grid, double *spec) {
  return guard([&] {
    NEED(s && grid && spec);
    std::vector<double2> out(4 * static_cast<size_t>(s->core->S));
    s->core->analyze_raw(PREP_SCALAR, grid, nullptr, nullptr, out.data());
    std::vector<double2> o4(out.begin() + 3 * s->core->S, out.end());
    s->view().from_packed(o4, spec);
  });
}
lmw_status lmw_sphere_analyze_divergence(lmw_sphere *s, const double *A, const double *B, double *spec) {
  return guard([&] {
    NEED(s && A && B && spec);
    std::vector<double2> out(4 * static_cast<size_t>(s->core->S));
    s->core->analyze_raw(PREP_DIV, A, B, nullptr, out.data());
    std::vector<double2> o1(out.begin(), out.begin() + s->core->S);
    s->view().from_packed(o1, spec);
  });
}
lmw_status lmw_sphere_tendencies(lmw_sphere *s, double *dzeta, double *ddelta, double *dphi) {
  return guard([&] {
    NEED(s && dzeta && ddelta && dphi);
    const int S = s->core->S;
    std::vector<double2> a(S), b(S), c(S);
    s->core->tendencies_raw(0, a.data(), b.data(), c.data());
    s->view().from_packed(a, dzeta);
    s->view().from_packed(b, ddelta);
    s->view().from_packed(c, dphi);
  });
}
lmw_status lmw_sphere_gauss(lmw_sphere *s, double *mu, double *w) {
  return guard([&] {
    NEED(s && mu && w);
    s->core->mu.download(mu, s->core->J, 0, s->core->stream);
    s->core->w.download(w, s->core->J, 0, s->core->stream);
    s->core->sync();
  });
}
lmw_status lmw_sphere_get_tables(lmw_sphere *s, double *P, double *H) {
  return guard([&] {
    NEED(s && P && H);
    SphereCore &c = *s->core;
    std::vector<double> ps(c.Ps.n), hs(c.Hs.n);
    c.Ps.download(ps.data(), ps.size(), 0, c.stream);
    c.Hs.download(hs.data(), hs.size(), 0, c.stream);
    c.sync();
    const int N = c.N;
    const size_t per = static_cast<size_t>(N + 1) * (N + 1);
    std::fill(P, P + per * c.J, 0.0);
    std::fill(H, H + per * c.J, 0.0);
    for (int jh = 0; jh < c.Jh; jh++)
      for (int m = 0; m <= N; m++)
        for (int k = 0; k <= N - m; k++) {
          const int r = (k & 1) ? c.h_kpe[m] + (k >> 1) : (k >> 1);
          const size_t src = (static_cast<size_t>(c.h_poff[m]) + r) * c.Jhp + jh;
          const double sgn = (k & 1) ? -1.0 : 1.0;  // P(-mu) = (-1)^(n-m) P(mu), H(-mu) = -(-1)^(n-m) H(mu)
          const int jn = c.Jh + jh, js = c.Jh - 1 - jh;
          P[(static_cast<size_t>(jn) * (N + 1) + m) * (N + 1) + k] = ps[src];
          H[(static_cast<size_t>(jn) * (N + 1) + m) * (N + 1) + k] = hs[src];
          P[(static_cast<size_t>(js) * (N + 1) + m) * (N + 1) + k] = sgn * ps[src];
          H[(static_cast<size_t>(js) * (N + 1) + m) * (N + 1) + k] = -sgn * hs[src];
        }
  });
}

// ------------------------------------------------------------------------------------------------ ensembles
lmw_status lmw_ensemble_create(int members, const int *truncations, const lmw_sphere_params *params, int precision, lmw_ensemble **out) {
  return guard([&] {
    NEED(out && truncations);
    *out = nullptr;
    if (members < 1) invalid("an ensemble needs at least one member");
    lmw_sphere_params p = params ? *params : lmw_sphere_params{6.37122e6, 7.292e-5, 9.80616, 0.0};
    std::map<int, std::vector<int>> by_t;
    for (int i = 0; i < members; i++) {
      if (truncations[i] < 4 || truncations[i] > 170) invalid("truncation must be between 4 and 170");
      by_t[truncations[i]].push_back(i);
    }
    auto e = std::make_unique<lmw_ensemble>();
    e->where.resize(members);
    for (auto &kv : by_t) {
      const int gi = static_cast<int>(e->groups.size());
      e->groups.push_back(std::make_unique<SphereCore>(kv.first, p, static_cast<int>(kv.second.size()), precision));
      for (size_t l = 0; l < kv.second.size(); l++) e->where[kv.second[l]] = {gi, static_cast<int>(l)};
    }
    *out = e.release();
  });
}
lmw_status lmw_ensemble_destroy(lmw_ensemble *e) {
  return guard([&] { delete e; });
}
lmw_status lmw_ensemble_member_dims(const lmw_ensemble *e, int member, int *truncation, int *nlat, int *nlon) {
  return guard([&] {
    NEED(e);
    MemberView v = const_cast<lmw_ensemble *>(e)->view(member);
    if (truncation) *truncation = v.c->N;
    if (nlat) *nlat = v.c->J;
    if (nlon) *nlon = v.c->K;
  });
}
lmw_status lmw_ensemble_set_state_grid(lmw_ensemble *e, int member, const double *h, const double *u, const double *v, size_t n) {
  return guard([&] {
    NEED(e);
    MemberView mv = e->view(member);
    if (n != mv.ngrid()) invalid("grid size mismatch");
    NEED(h && u && v);
    mv.c->set_state_grid(mv.b, h, u, v);
    mv.c->sync();
  });
}
}  // extern "C"

// The members [member0, member0 + count) as runs that sit in one group with consecutive local indices:
// fn(core, first local index, run length, offset of the run in the range) once per run.  same_truncation: refuse a
// range whose members differ in truncation (the host grids of one call share a size).
template <class F> static void member_runs(lmw_ensemble *e, int member0, int count, bool same_truncation, F &&fn) {
  if (count < 0) invalid("member count must not be negative");
  if (count == 0) return;
  if (member0 < 0 || member0 + static_cast<int64_t>(count) > static_cast<int64_t>(e->where.size())) invalid("member index out of range");
  const int N = e->view(member0).c->N;
  int i = member0;
  while (i < member0 + count) {
    MemberView v = e->view(i);
    if (same_truncation && v.c->N != N) invalid("the members of one call must share a truncation");
    int n = 1;
    while (i + n < member0 + count && e->where[i + n].first == e->where[i].first && e->where[i + n].second == v.b + n) n++;
    fn(*v.c, v.b, n, i - member0);
    i += n;
  }
}

extern "C" {
lmw_status lmw_ensemble_set_state_grids(lmw_ensemble *e, int member0, int count, const double *h, const double *u, const double *v, size_t n) {
  return guard([&] {
    NEED(e);
    if (count > 0) {
      MemberView mv = e->view(member0);
      if (n != static_cast<size_t>(count) * mv.ngrid()) invalid("grid size mismatch");
      NEED(h && u && v);
    }
    member_runs(e, member0, count, true, [&](SphereCore &c, int b0, int nb, int off) {
      const size_t o = static_cast<size_t>(off) * c.J * c.K;
      c.set_state_grids(b0, nb, h + o, u + o, v + o);
    });
  });
}
lmw_status lmw_ensemble_setup(lmw_ensemble *e, int member0, int count, int which, double alpha) {
  return guard([&] {
    NEED(e);
    if (which == LMW_SETUP_WILLIAMSON1) invalid("a tracer needs a single-member sphere");
    member_runs(e, member0, count, false, [&](SphereCore &c, int b0, int nb, int) { c.setup(b0, nb, which, alpha); });
  });
}
lmw_status lmw_ensemble_fields(lmw_ensemble *e, int member0, int count, const char *h, const char *u, const char *v) {
  return guard([&] {
    NEED(e);
    NEED(h);
    NEED(u);
    NEED(v);
    member_runs(e, member0, count, false, [&](SphereCore &c, int b0, int nb, int) { c.fields_expr(b0, nb, h, u, v); });
  });
}
lmw_status lmw_ensemble_orography(lmw_ensemble *e, int member0, int count, const char *hs) {
  return guard([&] {
    NEED(e);
    NEED(hs);
    member_runs(e, member0, count, false, [&](SphereCore &c, int b0, int nb, int) { c.orography_expr(b0, nb, hs); });
  });
}
lmw_status lmw_ensemble_set_orography_grid(lmw_ensemble *e, int member, const double *hs, size_t n) {
  return guard([&] {
    NEED(e);
    MemberView mv = e->view(member);
    if (n != mv.ngrid()) invalid("grid size mismatch");
    NEED(hs);
    mv.c->set_orography(mv.b, hs);
    mv.c->sync();
  });
}
lmw_status lmw_ensemble_set_axis_tilt(lmw_ensemble *e, int member, double alpha) {
  return guard([&] {
    NEED(e);
    MemberView mv = e->view(member);
    mv.c->set_axis_tilt(mv.b, alpha);
    mv.c->sync();
  });
}
lmw_status lmw_ensemble_set_spectral(lmw_ensemble *e, int member, int field, const double *re_im, size_t n_complex) {
  return guard([&] {
    NEED(e);
    MemberView mv = e->view(member);
    if (field < 0 || field > 2) invalid("unknown spectral field");
    if (n_complex != mv.nfull()) invalid("spectral size mismatch");
    NEED(re_im);
    std::vector<double2> p;
    mv.to_packed(re_im, p);
    mv.set_spectral(field, p);
    mv.c->sync();
  });
}
lmw_status lmw_ensemble_get_spectral(lmw_ensemble *e, int member, int field, double *re_im, size_t n_complex) {
  return guard([&] {
    NEED(e);
    MemberView mv = e->view(member);
    if (field < 0 || field > 2) invalid("unknown spectral field");
    if (n_complex != mv.nfull()) invalid("spectral size mismatch");
    NEED(re_im);
    mv.from_packed(mv.spectral(field), re_im);
  });
}
lmw_status lmw_ensemble_get_grid(lmw_ensemble *e, int member, int field, double *out, size_t n) {
  return guard([&] {
    NEED(e);
    MemberView mv = e->view(member);
    if (n != mv.ngrid()) invalid("grid size mismatch");
    NEED(out);
    mv.grid(field, out);
  });
}
lmw_status lmw_ensemble_diag(lmw_ensemble *e, int member, int which, double *out) {
  return guard([&] {
    NEED(e && out);
    *out = e->view(member).diag(which);
  });
}
lmw_status lmw_ensemble_groups(const lmw_ensemble *e, int *ngroups, int *truncations, int *members_per_group) {
  return guard([&] {
    NEED(e && ngroups);
    *ngroups = static_cast<int>(e->groups.size());
    for (size_t g = 0; g < e->groups.size(); g++) {
      if (truncations) truncations[g] = e->groups[g]->N;
      if (members_per_group) members_per_group[g] = e->groups[g]->B;
    }
  });
}
static void ensemble_steps(lmw_ensemble *e, const double *dts, const int64_t *counts, bool use_graph) {
  NEED(e && dts && counts);
  for (size_t g = 0; g < e->groups.size(); g++)
    if (counts[g] < 0) invalid("step count must not be negative");
  // each group runs on its own stream, so the truncations overlap on the device
  for (size_t g = 0; g < e->groups.size(); g++) {
    SphereCore &c = *e->groups[g];
    if (use_graph) c.steps_graph(dts[g], counts[g]);
    else c.steps(dts[g], counts[g]);
  }
  for (size_t g = 0; g < e->groups.size(); g++) {
    SphereCore &c = *e->groups[g];
    c.sync();
    for (int b = 0; b < c.B; b++)
      for (int64_t i = 0; i < counts[g]; i++) c.t[b] += dts[g];
  }
}
lmw_status lmw_ensemble_steps(lmw_ensemble *e, const double *dts, const int64_t *counts) {
  return guard([&] { ensemble_steps(e, dts, counts, false); });
}
lmw_status lmw_ensemble_steps_graph(lmw_ensemble *e, const double *dts, const int64_t *counts) {
  return guard([&] { ensemble_steps(e, dts, counts, true); });
}
// A graph of nmacro macro steps: group g advances nmacro x substeps[g] steps of dts[g] as one chain on its stream
// (its member parts forked once), the groups' chains concurrent branches.  The groups are independent models, so a
// graph of several macro steps lets the small groups' next macro steps run beside the largest group's last steps
// instead of waiting at every macro step's join; each group's state advances through the same steps in the same
// order either way.
static void capture_macro_graph(lmw_ensemble *e, const double *dts, const int *substeps, int nmacro, hipGraphExec_t &ge) {
  const size_t ng = e->groups.size();
  if (!e->origin) {
    LMW_HIP(hipStreamCreate(&e->origin));
    e->events.resize(ng + 1);
    for (auto &ev : e->events) LMW_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  if (ge) {
    LMW_HIP(hipGraphExecDestroy(ge));
    ge = nullptr;
  }
  hipGraph_t g;
  LMW_HIP(hipStreamBeginCapture(e->origin, hipStreamCaptureModeThreadLocal));
  LMW_HIP(hipEventRecord(e->events[0], e->origin));
  for (size_t i = 0; i < ng; i++) {
    SphereCore &c = *e->groups[i];
    LMW_HIP(hipStreamWaitEvent(c.stream, e->events[0], 0));
    c.enqueue_steps(dts[i], substeps[i] * nmacro);  // the group's member parts forked once for all its steps
    LMW_HIP(hipEventRecord(e->events[i + 1], c.stream));
    LMW_HIP(hipStreamWaitEvent(e->origin, e->events[i + 1], 0));
  }
  LMW_HIP(hipStreamEndCapture(e->origin, &g));
  LMW_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
  LMW_HIP(hipGraphDestroy(g));
}
static void capture_macro(lmw_ensemble *e, const double *dts, const int *substeps) {
  drop_macro(e);
  capture_macro_graph(e, dts, substeps, 1, e->macro);
  e->macro_dts.assign(dts, dts + e->groups.size());
  e->macro_sub.assign(substeps, substeps + e->groups.size());
}
// Persistent variant of the macro step: ticks of the finest group's step; at each tick the groups due a step run
// their four stages together, each stage as one multi-group inverse-transform launch, the groups' grid-point
// kernels, and one multi-group forward-transform launch, all on the origin stream.
static void capture_macro_persistent(lmw_ensemble *e, const double *dts, const int *substeps) {
  const size_t ng = e->groups.size();
  if (ng > static_cast<size_t>(kMaxGroups)) invalid("persistent macro steps serve at most 4 truncation groups");
  int T = 0;
  for (size_t i = 0; i < ng; i++) {
    if (e->groups[i]->precision == LMW_FP64) invalid("persistent macro steps need a mixed-precision ensemble");
    if (substeps[i] < 1) invalid("persistent macro steps need at least one sub-step per group");
    T = std::max(T, substeps[i]);
  }
  for (size_t i = 0; i < ng; i++)
    if (T % substeps[i]) invalid("persistent macro steps need sub-step counts that divide the largest one");
  std::vector<SphereCore *> gs = all_groups(e);
  std::vector<double> dtv(dts, dts + ng);
  std::vector<unsigned> masks(T);
  for (int t = 0; t < T; t++) {
    unsigned mask = 0;
    for (size_t i = 0; i < ng; i++)
      if (t % (T / substeps[i]) == 0) mask |= 1u << i;
    masks[t] = mask;
    auto &w = e->work[mask];
    if (!w) {
      w = std::make_unique<MultiWork>();
      build_multi_work(gs, mask, *w);
    }
  }
  if (!e->origin) {
    LMW_HIP(hipStreamCreate(&e->origin));
    e->events.resize(ng + 1);
    for (auto &ev : e->events) LMW_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  if (e->macro) {
    LMW_HIP(hipGraphExecDestroy(e->macro));
    e->macro = nullptr;
  }
  for (auto *c : gs) c->sync();
  hipGraph_t g;
  LMW_HIP(hipStreamBeginCapture(e->origin, hipStreamCaptureModeThreadLocal));
  for (int t = 0; t < T; t++) {
    const MultiWork &w = *e->work[masks[t]];
    for (int stage = 1; stage <= 4; stage++) {
      launch_synth_mx_multi(gs, w, e->origin);
      // the groups' grid-point kernels (different FFT lengths) as concurrent branches on their own streams
      LMW_HIP(hipEventRecord(e->events[0], e->origin));
      for (size_t i = 0; i < ng; i++) {
        if (!(masks[t] & (1u << i))) continue;
        LMW_HIP(hipStreamWaitEvent(gs[i]->stream, e->events[0], 0));
        gs[i]->launch_tend_grid(0, gs[i]->B, 1);
        LMW_HIP(hipEventRecord(e->events[i + 1], gs[i]->stream));
        LMW_HIP(hipStreamWaitEvent(e->origin, e->events[i + 1], 0));
      }
      launch_anal_mx_multi(gs, dtv, stage, w, e->origin);
    }
  }
  LMW_HIP(hipStreamEndCapture(e->origin, &g));
  LMW_HIP(hipGraphInstantiate(&e->macro, g, nullptr, nullptr, 0));
  LMW_HIP(hipGraphDestroy(g));
  e->macro_dts.assign(dts, dts + ng);
  e->macro_sub.assign(substeps, substeps + ng);
}
lmw_status lmw_ensemble_set_persistent(lmw_ensemble *e, int on) {
  return guard([&] {
    NEED(e);
    if (e->persistent != (on != 0)) drop_macro(e);
    e->persistent = on != 0;
  });
}
lmw_status lmw_ensemble_macro_steps(lmw_ensemble *e, const double *dts, const int *substeps, int64_t count) {
  return guard([&] {
    NEED(e && dts && substeps);
    const size_t ng = e->groups.size();
    if (count < 0) invalid("step count must not be negative");
    for (size_t i = 0; i < ng; i++) {
      if (substeps[i] < 0) invalid("step count must not be negative");
      if (substeps[i] > 0 && !(dts[i] > 0)) invalid("dt must be positive");
      if (e->groups[i]->has_tracer) invalid("macro steps do not carry tracers");
    }
    if (count == 0) return;
    for (auto &c : e->groups) c->sync();
    bool same = e->macro && e->macro_sub.size() == ng;
    for (size_t i = 0; same && i < ng; i++) same = e->macro_dts[i] == dts[i] && e->macro_sub[i] == substeps[i];
    if (!same) {
      if (e->persistent) capture_macro_persistent(e, dts, substeps);
      else capture_macro(e, dts, substeps);
    }
    int64_t k = 0;
    if (!e->persistent && count >= lmw_ensemble::kMacroGraph) {  // graphs of kMacroGraph macro steps, then single ones
      if (!e->macro_many) capture_macro_graph(e, dts, substeps, lmw_ensemble::kMacroGraph, e->macro_many);
      for (; k + lmw_ensemble::kMacroGraph <= count; k += lmw_ensemble::kMacroGraph) LMW_HIP(hipGraphLaunch(e->macro_many, e->origin));
    }
    for (; k < count; k++) LMW_HIP(hipGraphLaunch(e->macro, e->origin));
    LMW_HIP(hipStreamSynchronize(e->origin));
    for (size_t i = 0; i < ng; i++) {
      SphereCore &c = *e->groups[i];
      for (int b = 0; b < c.B; b++)
        for (int64_t k = 0; k < count; k++)
          for (int q = 0; q < substeps[i]; q++) c.t[b] += dts[i];
    }
  });
}
lmw_status lmw_ensemble_steps_timed(lmw_ensemble *e, const double *dts, const int64_t *counts, double *kernel_ms) {
  return guard([&] {
    NEED(e && dts && counts && kernel_ms);
    for (size_t g = 0; g < e->groups.size(); g++) {
      if (counts[g] < 0) invalid("step count must not be negative");
      SphereCore &c = *e->groups[g];
      double *ms = kernel_ms + 3 * g;
      ms[0] = ms[1] = ms[2] = 0;
      c.steps_timed(dts[g], counts[g], ms);  // groups one after another: each kernel is timed alone on the device
      for (int b = 0; b < c.B; b++)
        for (int64_t i = 0; i < counts[g]; i++) c.t[b] += dts[g];
    }
  });
}

}  // extern "C"

namespace lmw {
// the device state behind the handles, for the corrector's entry points (corrector.hip)
SphereCore &core_of(lmw_sphere *s) {
  if (!s) invalid("null pointer argument");
  return *s->core;
}
std::vector<SphereCore *> all_groups(lmw_ensemble *e) {
  if (!e) invalid("null pointer argument");
  std::vector<SphereCore *> out;
  for (auto &g : e->groups) out.push_back(g.get());
  return out;
}
void drop_macro(lmw_ensemble *e) {
  if (e && e->macro) {
    LMW_HIP(hipGraphExecDestroy(e->macro));
    e->macro = nullptr;
  }
  if (e && e->macro_many) {
    LMW_HIP(hipGraphExecDestroy(e->macro_many));
    e->macro_many = nullptr;
  }
}
SphereCore &single_group(lmw_ensemble *e) {
  if (!e) invalid("null pointer argument");
  if (e->groups.size() != 1) invalid("the ensemble must hold members of a single truncation");
  return *e->groups[0];  // one group: local member index == member index
}
}  // namespace lmw
