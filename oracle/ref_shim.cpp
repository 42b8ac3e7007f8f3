// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C-linkage shim over the reference's own lmweather::Sphere / lmweather::Plane (cpp/sphere.hpp, cpp/plane.hpp),
// compiled together with the reference's cpp/sphere.cpp and cpp/plane.cpp (read in place, never copied) into
// oracle/_ref/liblmw_ref.so by oracle/Makefile.  It exposes the same C API as lmw_oracle.c with a ref_ prefix, so
// a test can drive both through one sequence of calls and compare them bit for bit.  The private state (spectral
// coefficients, tendencies) is reached by compiling the reference header with its private members opened up: that
// changes access checking only, not layout or code.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>
#define private public
#include "plane.hpp"
#include "sphere.hpp"
#undef private
#include "expr.hpp"

using lmweather::cplx;
using lmweather::Plane;
using lmweather::PlaneParams;
using lmweather::Sphere;
using lmweather::SphereParams;

namespace {
thread_local std::string g_err;
template <class F> int guard(F &&fn) {
  try { fn(); return 0; }
  catch (const std::invalid_argument &e) { g_err = e.what(); return 1; }
  catch (const std::exception &e) { g_err = e.what(); return 2; }
}
std::vector<double> vec(const double *p, size_t n) { return std::vector<double>(p, p + n); }
void put(const std::vector<double> &v, double *out) { std::memcpy(out, v.data(), v.size() * sizeof(double)); }
void putc(const std::vector<cplx> &v, double *out) { std::memcpy(out, v.data(), v.size() * sizeof(cplx)); }
std::vector<cplx> getc(const double *p, size_t n) { return std::vector<cplx>(reinterpret_cast<const cplx *>(p), reinterpret_cast<const cplx *>(p) + n); }
}  // namespace

extern "C" {
const char *ref_last_error(void) { return g_err.c_str(); }
// the reference's expression compiler (cpp/expr.hpp) over two variables and named constants: out[i] = e(x0[i], x1[i])
int ref_expr_eval(const char *text, const char *v0, const char *v1, int np, const char *const *pn, const double *pv, const double *x0,
                  const double *x1, size_t n, double *out) {
  return guard([&] {
    std::map<std::string, double> par;
    for (int i = 0; i < np; i++) par[pn[i]] = pv[i];
    lmnum::Expr e(text, {v0, v1}, par);
    for (size_t i = 0; i < n; i++) {
      const double x[2] = {x0[i], x1[i]};
      out[i] = e.eval(x);
    }
  });
}

int ref_fft(double *re_im, size_t n, int sign) {
  return guard([&] {
    std::vector<cplx> a = getc(re_im, n);
    lmweather::fft(a, sign);
    putc(a, re_im);
  });
}
void ref_gauss_legendre(int n, double *x, double *w) {
  std::vector<double> xx, ww;
  lmweather::gauss_legendre(n, xx, ww);
  put(xx, x);
  put(ww, w);
}
void ref_legendre_column(int m, int nmax, double mu, double *out) {
  std::vector<double> c;
  lmweather::legendre_column(m, nmax, mu, c);
  put(c, out);
}

// ------------------------------------------------------------------------------------------------ sphere
struct RefSphere { Sphere s; RefSphere(int n, const SphereParams &p) : s(n, p) {} };
void *ref_sphere_create(int N, double radius, double omega, double gravity, double diffusion) {
  RefSphere *out = nullptr;
  SphereParams p;
  p.radius = radius; p.omega = omega; p.gravity = gravity; p.diffusion = diffusion;
  if (guard([&] { out = new RefSphere(N, p); })) return nullptr;
  return out;
}
void ref_sphere_destroy(void *h) { delete static_cast<RefSphere *>(h); }
#define S (static_cast<RefSphere *>(h)->s)
#define CS (static_cast<const RefSphere *>(h)->s)
void ref_sphere_dims(const void *h, int *nlat, int *nlon) { *nlat = CS.nlat(); *nlon = CS.nlon(); }
double ref_sphere_time(const void *h) { return CS.time(); }
void ref_sphere_gauss(const void *h, double *mu, double *w) { put(CS.gauss_mu(), mu); put(CS.gauss_weights(), w); }
void ref_sphere_set_diffusion(void *h, double nu4) { S.set_diffusion(nu4); }
void ref_sphere_set_advect_only(void *h, int on) { S.set_advect_only(on != 0); }
void ref_sphere_set_axis_tilt(void *h, double a) { S.set_axis_tilt(a); }
int ref_sphere_set_orography_grid(void *h, const double *hs, size_t n) { return guard([&] { S.set_orography_grid(vec(hs, n)); }); }
void ref_sphere_mark_reference(void *h) { S.mark_reference(); }
int ref_sphere_set_state_grid(void *h, const double *hh, const double *u, const double *v, size_t n) {
  return guard([&] { S.set_state_grid(vec(hh, n), vec(u, n), vec(v, n)); });
}
int ref_sphere_set_tracer_grid(void *h, const double *q, size_t n) { return guard([&] { S.set_tracer_grid(vec(q, n)); }); }
int ref_sphere_add_mode(void *h, int n, int m, double amp, double depth) { return guard([&] { S.add_mode(n, m, amp, depth); }); }
int ref_sphere_perturb(void *h, double amp, uint64_t seed, int max_wave) { return guard([&] { S.perturb(amp, seed, max_wave); }); }
void ref_sphere_step(void *h, double dt) { S.step(dt); }
int ref_sphere_run(void *h, double seconds, double dt) { return guard([&] { S.run(seconds, dt); }); }
int ref_sphere_stable_dt(const void *h, double safety, double *out) { return guard([&] { *out = CS.stable_dt(safety); }); }
void ref_sphere_grid_height(const void *h, double *g) { put(CS.grid_height(), g); }
void ref_sphere_grid_vorticity(const void *h, double *g) { put(CS.grid_vorticity(), g); }
void ref_sphere_grid_u(const void *h, double *g) { put(CS.grid_u(), g); }
void ref_sphere_grid_v(const void *h, double *g) { put(CS.grid_v(), g); }
int ref_sphere_grid_tracer(const void *h, double *g) { return guard([&] { put(CS.grid_tracer(), g); }); }
double ref_sphere_mean_height(const void *h) { return CS.mean_height(); }
double ref_sphere_kinetic_energy(const void *h) { return CS.kinetic_energy(); }
double ref_sphere_potential_energy(const void *h) { return CS.potential_energy(); }
double ref_sphere_energy(const void *h) { return CS.energy(); }
void ref_sphere_spectrum(const void *h, double *ke, double *pe) {
  std::vector<double> a, b;
  CS.spectrum(a, b);
  put(a, ke);
  put(b, pe);
}
double ref_sphere_potential_enstrophy(const void *h) { return CS.potential_enstrophy(); }
double ref_sphere_max_wind(const void *h) { return CS.max_wind(); }
void ref_sphere_height_errors(const void *h, double *l2, double *linf) { *l2 = CS.height_error_l2(); *linf = CS.height_error_linf(); }
int ref_sphere_tracer_errors(const void *h, double *l2, double *linf) { return guard([&] { *l2 = CS.tracer_error_l2(); *linf = CS.tracer_error_linf(); }); }
int ref_sphere_tracer_mean(const void *h, double *out) { return guard([&] { *out = CS.tracer_mean(); }); }
int ref_sphere_tracer_mass(const void *h, double *out) { return guard([&] { *out = CS.tracer_mass(); }); }
double ref_sphere_height_at(const void *h, double lat, double lon) { return CS.height_at(lat, lon); }
double ref_sphere_vorticity_at(const void *h, double lat, double lon) { return CS.vorticity_at(lat, lon); }
int ref_sphere_tracer_at(const void *h, double lat, double lon, double *out) { return guard([&] { *out = CS.tracer_at(lat, lon); }); }
double ref_sphere_wind_at(const void *h, double lat, double lon, int zonal) { return CS.wind_at(lat, lon, zonal != 0); }
int ref_sphere_mode_coefficient(const void *h, int n, int m, int imag, double *out) { return guard([&] { *out = CS.mode_coefficient(n, m, imag != 0); }); }
int ref_sphere_wave_crest(const void *h, int m, double lat, double *out) { return guard([&] { *out = CS.wave_crest(m, lat); }); }
void ref_sphere_synthesize(const void *h, const double *spec, double *grid) {
  std::vector<double> g;
  size_t ns = static_cast<size_t>(CS.truncation() + 1) * (CS.truncation() + 1);
  CS.synthesize_scalar(getc(spec, ns), g);
  put(g, grid);
}
void ref_sphere_analyze(const void *h, const double *grid, double *spec) {
  std::vector<cplx> sp;
  CS.analyze_scalar(vec(grid, static_cast<size_t>(CS.nlat()) * CS.nlon()), sp);
  putc(sp, spec);
}
void ref_sphere_analyze_divergence(const void *h, const double *A, const double *B, double *spec) {
  std::vector<cplx> sp;
  size_t ng = static_cast<size_t>(CS.nlat()) * CS.nlon();
  CS.analyze_divergence(vec(A, ng), vec(B, ng), sp);
  putc(sp, spec);
}
void ref_sphere_winds(const void *h, const double *z, const double *d, double *U, double *V) {
  std::vector<double> u, v;
  size_t ns = static_cast<size_t>(CS.truncation() + 1) * (CS.truncation() + 1);
  CS.winds_spectral_to_grid(getc(z, ns), getc(d, ns), u, v);
  put(u, U);
  put(v, V);
}
void ref_sphere_tendencies(const void *h, double *dz, double *dd, double *df) {
  std::vector<cplx> a, b, c;
  CS.tendencies(CS.zeta_, CS.delta_, CS.phi_, a, b, c);
  putc(a, dz);
  putc(b, dd);
  putc(c, df);
}
void ref_sphere_get_spectral(void *h, int field, double *out) {
  putc(field == 0 ? S.zeta_ : field == 1 ? S.delta_ : field == 2 ? S.phi_ : S.q_, out);
}
void ref_sphere_set_spectral(void *h, int field, const double *in) {
  std::vector<cplx> v = getc(in, S.zeta_.size());
  if (field == 0) S.zeta_ = v;
  else if (field == 1) S.delta_ = v;
  else if (field == 2) S.phi_ = v;
  else { S.q_ = v; S.has_tracer_ = true; }
}
void ref_sphere_get_tables(const void *h, double *P, double *H) {
  const int N = CS.truncation();
  for (int j = 0; j < CS.nlat(); j++)
    for (int m = 0; m <= N; m++)
      for (int k = 0; k <= N; k++) {
        size_t o = (static_cast<size_t>(j) * (N + 1) + m) * (N + 1) + k;
        P[o] = k < N - m + 1 ? CS.P_[j][m][k] : 0.0;
        H[o] = k < N - m + 1 ? CS.H_[j][m][k] : 0.0;
      }
}
void ref_sphere_get_coriolis(const void *h, double *f) { put(CS.f_grid_, f); }
#undef S
#undef CS

// ------------------------------------------------------------------------------------------------ plane
void *ref_plane_create(int nx, int ny, double lx, double ly, double gravity, double depth, double f0, double beta, double viscosity, double drag,
                       int linear) {
  PlaneParams p;
  p.nx = nx; p.ny = ny; p.lx = lx; p.ly = ly; p.gravity = gravity; p.depth = depth; p.f0 = f0; p.beta = beta;
  p.viscosity = viscosity; p.drag = drag; p.linear = linear != 0;
  Plane *out = nullptr;
  if (guard([&] { out = new Plane(p); })) return nullptr;
  return out;
}
void ref_plane_destroy(void *h) { delete static_cast<Plane *>(h); }
#define P (*static_cast<Plane *>(h))
#define CP (*static_cast<const Plane *>(h))
void ref_plane_set_params(void *h, double gravity, double depth, double f0, double beta, double viscosity, double drag, int linear) {
  PlaneParams &p = P.params();
  p.gravity = gravity; p.depth = depth; p.f0 = f0; p.beta = beta; p.viscosity = viscosity; p.drag = drag; p.linear = linear != 0;
}
double ref_plane_time(const void *h) { return CP.time(); }
void ref_plane_reset(void *h) { P.reset(); }
void ref_plane_perturb(void *h, double amp, uint64_t seed) { P.perturb(amp, seed); }
void ref_plane_step(void *h, double dt) { P.step(dt); }
int ref_plane_run(void *h, double seconds, double dt) { return guard([&] { P.run(seconds, dt); }); }
int ref_plane_stable_dt(const void *h, double safety, double *out) { return guard([&] { *out = CP.stable_dt(safety); }); }
double ref_plane_mean_depth(const void *h) { return CP.mean_depth(); }
double ref_plane_energy(const void *h) { return CP.energy(); }
double ref_plane_max_speed(const void *h) { return CP.max_speed(); }
double ref_plane_pv_anomaly_rms(const void *h) { return CP.potential_vorticity_anomaly_rms(); }
double ref_plane_zeta_at_corner(const void *h, int i, int j) { return CP.zeta_at_corner(i, j); }
double ref_plane_at(const void *h, double x, double y, int kind) { return kind == 0 ? CP.h_at(x, y) : kind == 1 ? CP.u_at(x, y) : CP.v_at(x, y); }
int ref_plane_zonal_amplitude(const void *h, int m, int j, double *out) { return guard([&] { *out = CP.zonal_amplitude(m, j); }); }
int ref_plane_zonal_crest(const void *h, int m, int j, double *out) { return guard([&] { *out = CP.zonal_crest(m, j); }); }
int ref_plane_harmonic_energy(const void *h, int m, double *out) { return guard([&] { *out = CP.harmonic_energy(m); }); }
int ref_plane_v_zonal_crest(const void *h, int m, int j, double *out) { return guard([&] { *out = CP.v_zonal_crest(m, j); }); }
int ref_plane_set_field(void *h, int field, const double *g, size_t n) {
  return guard([&] {
    if (field == 0) P.set_h_from_grid(vec(g, n));
    else if (field == 1) P.set_u_from_grid(vec(g, n));
    else P.set_v_from_grid(vec(g, n));
  });
}
void ref_plane_get_field(const void *h, int field, double *out) { put(field == 0 ? CP.h_ : field == 1 ? CP.u_ : CP.v_, out); }
void ref_plane_tendency(const void *h, double *dh, double *du, double *dv) {
  std::vector<double> a, b, c;
  CP.tendency(CP.h_, CP.u_, CP.v_, a, b, c);
  put(a, dh);
  put(b, du);
  put(c, dv);
}
#undef P
#undef CP
}
