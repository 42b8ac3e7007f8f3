// lmw.hpp -- header-only C++ façade over the C ABI (lmw.h) with the reference's class interface: lmw::Sphere
// mirrors lmweather::Sphere (cpp/sphere.hpp:27-121) and lmw::Plane mirrors lmweather::Plane (cpp/plane.hpp:20-68).
// Errors come back as std::invalid_argument / std::runtime_error carrying the reference's exact texts; a device
// failure is a std::runtime_error as well.  A host that compiled against the reference classes switches by
// replacing `lmweather::` with `lmw::` and linking liblmw_hip.so.
#pragma once
#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "lmw.h"

namespace lmw {

using cplx = std::complex<double>;

inline void check(lmw_status s) {
  if (s == LMW_OK) return;
  const std::string msg = lmw_last_error();
  if (s == LMW_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

// fft(std::vector<cplx>&, int sign) and gauss_legendre(n, x, w) of cpp/sphere.hpp:14-16, on the GPU
inline void fft(std::vector<cplx> &a, int sign) { check(lmw_fft(reinterpret_cast<double *>(a.data()), a.size(), 1, sign)); }
inline void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  check(lmw_gauss_legendre(n, x.data(), w.data()));
}

struct SphereParams {  // cpp/sphere.hpp:20-25
  double radius = 6.37122e6;
  double omega = 7.292e-5;
  double gravity = 9.80616;
  double diffusion = 0;
};

class Sphere {
 public:
  Sphere(int truncation, const SphereParams &p) : p_(p) {
    lmw_sphere_params q{p.radius, p.omega, p.gravity, p.diffusion};
    check(lmw_sphere_create(truncation, &q, &h_));
    check(lmw_sphere_dims(h_, &N_, &nlat_, &nlon_));
  }
  ~Sphere() { lmw_sphere_destroy(h_); }
  Sphere(const Sphere &) = delete;
  Sphere &operator=(const Sphere &) = delete;

  int truncation() const { return N_; }
  int nlat() const { return nlat_; }
  int nlon() const { return nlon_; }
  const SphereParams &params() const { return p_; }
  double time() const { return diag(LMW_DIAG_TIME); }
  double lat(int j) const { return std::asin(gauss_mu()[j]); }
  double lon(int k) const { return 2 * 3.14159265358979323846 * k / nlon_; }

  void set_state_grid(const std::vector<double> &h, const std::vector<double> &u, const std::vector<double> &v) {
    const size_t n = (h.size() == u.size() && u.size() == v.size()) ? h.size() : 0;
    check(lmw_sphere_set_state_grid(h_, h.data(), u.data(), v.data(), n));
  }
  void set_axis_tilt(double alpha) { check(lmw_sphere_set_axis_tilt(h_, alpha)); }
  void set_diffusion(double nu4) {
    p_.diffusion = nu4;
    check(lmw_sphere_set_diffusion(h_, nu4));
  }
  void mark_reference() { check(lmw_sphere_mark_reference(h_)); }
  void add_mode(int n, int m, double amplitude, double mean_depth) { check(lmw_sphere_add_mode(h_, n, m, amplitude, mean_depth)); }
  void perturb(double amplitude, uint64_t seed, int max_wave) { check(lmw_sphere_perturb(h_, amplitude, seed, max_wave)); }
  void set_orography_grid(const std::vector<double> &hs) { check(lmw_sphere_set_orography_grid(h_, hs.data(), hs.size())); }
  void set_precision(int precision) { check(lmw_sphere_set_precision(h_, precision)); }
  // passive tracer (cpp/sphere.hpp:52-64)
  void set_tracer_grid(const std::vector<double> &q) { check(lmw_sphere_set_tracer_grid(h_, q.data(), q.size())); }
  void set_advect_only(bool on) { check(lmw_sphere_set_advect_only(h_, on ? 1 : 0)); }
  bool has_tracer() const {
    int t = 0;
    check(lmw_sphere_tracer_flags(h_, &t, nullptr));
    return t != 0;
  }
  bool advect_only() const {
    int a = 0;
    check(lmw_sphere_tracer_flags(h_, nullptr, &a));
    return a != 0;
  }
  std::vector<double> grid_tracer() const { return grid(LMW_GRID_TRACER); }
  double tracer_mean() const { return diag(LMW_DIAG_TRACER_MEAN); }
  double tracer_mass() const { return diag(LMW_DIAG_TRACER_MASS); }
  double tracer_at(double lat, double lon) const { return point(LMW_POINT_TRACER, lat, lon); }
  double tracer_min() const { return diag(LMW_DIAG_TRACER_MIN); }
  double tracer_max() const { return diag(LMW_DIAG_TRACER_MAX); }
  double tracer_error_l2() const { return diag(LMW_DIAG_TRACER_ERROR_L2); }
  double tracer_error_linf() const { return diag(LMW_DIAG_TRACER_ERROR_LINF); }

  void step(double dt) { check(lmw_sphere_step(h_, dt)); }
  void steps(double dt, int64_t count) { check(lmw_sphere_steps(h_, dt, count)); }
  void run(double seconds, double dt) { check(lmw_sphere_run(h_, seconds, dt)); }
  double stable_dt(double safety = 1.5) const {
    double out;
    check(lmw_sphere_stable_dt(h_, safety, &out));
    return out;
  }

  double mean_height() const { return diag(LMW_DIAG_MEAN_HEIGHT); }
  double energy() const { return diag(LMW_DIAG_ENERGY); }
  double kinetic_energy() const { return diag(LMW_DIAG_KINETIC); }
  double potential_energy() const { return diag(LMW_DIAG_POTENTIAL); }
  double potential_enstrophy() const { return diag(LMW_DIAG_ENSTROPHY); }
  double max_wind() const { return diag(LMW_DIAG_MAX_WIND); }
  double height_error_l2() const { return diag(LMW_DIAG_ERROR_L2); }
  double height_error_linf() const { return diag(LMW_DIAG_ERROR_LINF); }
  void spectrum(std::vector<double> &ke, std::vector<double> &pe) const {
    ke.assign(N_ + 1, 0.0);
    pe.assign(N_ + 1, 0.0);
    check(lmw_sphere_spectrum(h_, ke.data(), pe.data(), N_ + 1));
  }
  double height_at(double lat, double lon) const { return point(LMW_POINT_HEIGHT, lat, lon); }
  double wind_at(double lat, double lon, bool zonal) const { return point(zonal ? LMW_POINT_U : LMW_POINT_V, lat, lon); }
  double vorticity_at(double lat, double lon) const { return point(LMW_POINT_VORTICITY, lat, lon); }
  double mode_coefficient(int n, int m, bool imag) const {
    double out;
    check(lmw_sphere_mode_coefficient(h_, n, m, imag ? 1 : 0, &out));
    return out;
  }
  double wave_crest(int m, double lat) const {
    double out;
    check(lmw_sphere_wave_crest(h_, m, lat, &out));
    return out;
  }

  std::vector<double> grid_height() const { return grid(LMW_GRID_H); }
  std::vector<double> grid_u() const { return grid(LMW_GRID_U); }
  std::vector<double> grid_v() const { return grid(LMW_GRID_V); }
  std::vector<double> grid_vorticity() const { return grid(LMW_GRID_VORTICITY); }
  std::vector<double> grid_orography() const { return grid(LMW_GRID_OROGRAPHY); }
  // a Williamson et al. (1992) test case set up on the device (LMW_SETUP_*; the williamson1/2/5/6 operators)
  void setup(int which, double alpha = 0.0) { check(lmw_sphere_setup(h_, which, alpha)); }
  // the expression operators Sphere.fields / orography / tracer (cpp/main.cpp:158-172,212-228,252-268)
  void fields(const std::string &h, const std::string &u, const std::string &v) { check(lmw_sphere_fields(h_, h.c_str(), u.c_str(), v.c_str())); }
  void orography(const std::string &hs) { check(lmw_sphere_orography(h_, hs.c_str())); }
  void tracer(const std::string &q) { check(lmw_sphere_tracer(h_, q.c_str())); }
  // the text files of the Sphere.save_csv / save_spectrum / save_zonal_mean operators (cpp/main.cpp:328-383)
  void save_csv(const std::string &path) const { check(lmw_sphere_save_csv(h_, path.c_str())); }
  void save_spectrum(const std::string &path) const { check(lmw_sphere_save_spectrum(h_, path.c_str())); }
  void save_zonal_mean(const std::string &path) const { check(lmw_sphere_save_zonal_mean(h_, path.c_str())); }
  std::vector<double> gauss_mu() const {
    std::vector<double> mu(nlat_), w(nlat_);
    check(lmw_sphere_gauss(h_, mu.data(), w.data()));
    return mu;
  }
  std::vector<double> gauss_weights() const {
    std::vector<double> mu(nlat_), w(nlat_);
    check(lmw_sphere_gauss(h_, mu.data(), w.data()));
    return w;
  }

  // exposed for tests (cpp/sphere.hpp:97-100)
  void analyze_scalar(const std::vector<double> &grid, std::vector<cplx> &spec) const {
    spec.assign(static_cast<size_t>(N_ + 1) * (N_ + 1), 0);
    check(lmw_sphere_analyze(h_, grid.data(), reinterpret_cast<double *>(spec.data())));
  }
  void synthesize_scalar(const std::vector<cplx> &spec, std::vector<double> &grid) const {
    grid.assign(static_cast<size_t>(nlat_) * nlon_, 0.0);
    check(lmw_sphere_synthesize(h_, reinterpret_cast<const double *>(spec.data()), grid.data()));
  }
  void analyze_divergence(const std::vector<double> &A, const std::vector<double> &B, std::vector<cplx> &spec) const {
    spec.assign(static_cast<size_t>(N_ + 1) * (N_ + 1), 0);
    check(lmw_sphere_analyze_divergence(h_, A.data(), B.data(), reinterpret_cast<double *>(spec.data())));
  }

 private:
  lmw_sphere *h_ = nullptr;
  SphereParams p_;
  int N_ = 0, nlat_ = 0, nlon_ = 0;
  double diag(int which) const {
    double out;
    check(lmw_sphere_diag(h_, which, &out));
    return out;
  }
  double point(int which, double lat, double lon) const {
    double out;
    check(lmw_sphere_point(h_, which, lat, lon, &out));
    return out;
  }
  std::vector<double> grid(int field) const {
    std::vector<double> g(static_cast<size_t>(nlat_) * nlon_);
    check(lmw_sphere_get_grid(h_, field, g.data(), g.size()));
    return g;
  }
};

struct PlaneParams {  // cpp/plane.hpp:9-18
  int nx = 0, ny = 0;
  double lx = 0, ly = 0;
  double gravity = 9.81;
  double depth = 100;
  double f0 = 0, beta = 0;
  double viscosity = 0;
  double drag = 0;
  bool linear = false;
};

class Plane {
 public:
  explicit Plane(const PlaneParams &p, int precision = LMW_FP64) : p_(p) {
    lmw_plane_params q = raw(p);
    check(lmw_plane_create(&q, 1, precision, &h_));
    pull();
  }
  ~Plane() { lmw_plane_destroy(h_); }
  Plane(const Plane &) = delete;
  Plane &operator=(const Plane &) = delete;

  const PlaneParams &params() const { return p_; }
  // the parameter setters of the Plane operators (cpp/main.cpp:387-430); nx, ny, lx, ly stay fixed
  void set_params(const PlaneParams &p) {
    p_ = p;
    lmw_plane_params q = raw(p);
    check(lmw_plane_set_params(h_, &q));
  }
  double time() const {
    double out;
    check(lmw_plane_diag(h_, 0, LMW_PDIAG_TIME, &out));
    return out;
  }
  double dx() const { return p_.lx / p_.nx; }
  double dy() const { return p_.ly / p_.ny; }
  double x_center(int i) const { return (i + 0.5) * dx(); }
  double y_center(int j) const { return (j + 0.5) * dy(); }
  // field access on a host copy: edit with h(i, j) = ..., then push(); pull() refreshes after stepping
  double &h(int i, int j) { return h_host_[static_cast<size_t>(j) * p_.nx + i]; }
  double &u(int i, int j) { return u_host_[static_cast<size_t>(j) * p_.nx + i]; }
  double &v(int i, int j) { return v_host_[static_cast<size_t>(j) * p_.nx + i]; }
  void push() {
    check(lmw_plane_set_field(h_, 0, LMW_PLANE_H, h_host_.data(), h_host_.size()));
    check(lmw_plane_set_field(h_, 0, LMW_PLANE_U, u_host_.data(), u_host_.size()));
    check(lmw_plane_set_field(h_, 0, LMW_PLANE_V, v_host_.data(), v_host_.size()));
  }
  void pull() {
    h_host_.assign(static_cast<size_t>(p_.nx) * p_.ny, 0.0);
    u_host_.assign(h_host_.size(), 0.0);
    v_host_.assign(static_cast<size_t>(p_.nx) * (p_.ny + 1), 0.0);
    check(lmw_plane_get_field(h_, 0, LMW_PLANE_H, h_host_.data(), h_host_.size()));
    check(lmw_plane_get_field(h_, 0, LMW_PLANE_U, u_host_.data(), u_host_.size()));
    check(lmw_plane_get_field(h_, 0, LMW_PLANE_V, v_host_.data(), v_host_.size()));
  }
  void reset() {
    check(lmw_plane_reset(h_));
    pull();
  }
  void perturb(double amplitude, uint64_t seed) {
    check(lmw_plane_perturb(h_, 0, amplitude, seed));
    pull();
  }
  void step(double dt) {
    check(lmw_plane_step(h_, dt));
    pull();
  }
  void run(double seconds, double dt) {
    check(lmw_plane_run(h_, seconds, dt));
    pull();
  }
  double stable_dt(double safety = 0.5) const {
    double out;
    check(lmw_plane_stable_dt(h_, safety, &out));
    return out;
  }
  double mean_depth() const { return diag(LMW_PDIAG_MEAN_DEPTH); }
  double energy() const { return diag(LMW_PDIAG_ENERGY); }
  double max_speed() const { return diag(LMW_PDIAG_MAX_SPEED); }
  double potential_vorticity_anomaly_rms() const { return diag(LMW_PDIAG_PV_ANOMALY_RMS); }
  // bilinear in each field's own points (cpp/plane.cpp:245-263), on the host copy
  double h_at(double x, double y) const { return interp(x, y, 0); }
  double u_at(double x, double y) const { return interp(x, y, 1); }
  double v_at(double x, double y) const { return interp(x, y, 2); }
  // the Plane.wave_speed and Plane.rossby_radius operators (cpp/main.cpp:445-449), from the parameters
  double wave_speed() const { return std::sqrt(p_.gravity * p_.depth); }
  double rossby_radius() const {
    if (p_.f0 == 0) throw std::invalid_argument("Plane.rossby_radius: f0 is zero");
    return std::sqrt(p_.gravity * p_.depth) / std::fabs(p_.f0);
  }
  // zonal harmonics of the device state (cpp/plane.cpp:264-310)
  double zonal_amplitude(int m, int j) const { return harmonic(lmw_plane_zonal_amplitude, m, j); }
  double zonal_crest(int m, int j) const { return harmonic(lmw_plane_zonal_crest, m, j); }
  double v_zonal_crest(int m, int j) const { return harmonic(lmw_plane_v_zonal_crest, m, j); }
  double harmonic_energy(int m) const {
    double out;
    check(lmw_plane_harmonic_energy(h_, 0, m, &out));
    return out;
  }
  // cpp/plane.cpp:220-226, on the host copy
  double zeta_at_corner(int i, int j) const {
    const int nx = p_.nx, ny = p_.ny, im = (i + nx - 1) % nx;
    auto U = [&](int a, int b) { return u_host_[static_cast<size_t>(b) * nx + a]; };
    auto V = [&](int a, int b) { return v_host_[static_cast<size_t>(b) * nx + a]; };
    const double dvdx = (V(i % nx, j) - V(im, j)) / dx();
    const double uj = j < ny ? U(i % nx, j) : U(i % nx, ny - 1), ujm = j > 0 ? U(i % nx, j - 1) : U(i % nx, 0);
    return dvdx - (uj - ujm) / dy();
  }
  // the Plane.fields operator (cpp/main.cpp:410-431): h = H + h(x, y), u, v from expressions
  void fields(const std::string &h, const std::string &u, const std::string &v) {
    check(lmw_plane_fields(h_, 0, h.c_str(), u.c_str(), v.c_str()));
    pull();
  }
  // the text files of the Plane.save_csv / save_row / save_column operators (cpp/main.cpp:464-493)
  void save_csv(const std::string &path) const { check(lmw_plane_save_csv(h_, 0, path.c_str())); }
  void save_row(int row, const std::string &path) const { check(lmw_plane_save_row(h_, 0, row, path.c_str())); }
  void save_column(int column, const std::string &path) const { check(lmw_plane_save_column(h_, 0, column, path.c_str())); }
  void set_h_from_grid(const std::vector<double> &g) { set(LMW_PLANE_H, g); }
  void set_u_from_grid(const std::vector<double> &g) { set(LMW_PLANE_U, g); }
  void set_v_from_grid(const std::vector<double> &g) { set(LMW_PLANE_V, g); }

 private:
  lmw_plane *h_ = nullptr;
  PlaneParams p_;
  std::vector<double> h_host_, u_host_, v_host_;
  static lmw_plane_params raw(const PlaneParams &p) {
    return lmw_plane_params{p.nx, p.ny, p.lx, p.ly, p.gravity, p.depth, p.f0, p.beta, p.viscosity, p.drag, p.linear ? 1 : 0};
  }
  double diag(int which) const {
    double out;
    check(lmw_plane_diag(h_, 0, which, &out));
    return out;
  }
  double harmonic(lmw_status (*fn)(lmw_plane *, int, int, int, double *), int m, int j) const {
    double out;
    check(fn(h_, 0, m, j, &out));
    return out;
  }
  void set(int field, const std::vector<double> &g) {
    check(lmw_plane_set_field(h_, 0, field, g.data(), g.size()));
    pull();
  }
  double interp(double x, double y, int kind) const {
    const double dxx = p_.lx / p_.nx, dyy = p_.ly / p_.ny;
    const double xs = kind == 1 ? 0.0 : 0.5, ys = kind == 2 ? 0.0 : 0.5;
    const double fx = x / dxx - xs, fy = y / dyy - ys;
    const int rows = kind == 2 ? p_.ny + 1 : p_.ny;
    const double fl = std::floor(fx);
    const int i0 = static_cast<int>(fl), j0 = static_cast<int>(std::floor(fy));
    const double ax = fx - fl, ay = fy - std::floor(fy);
    const std::vector<double> &f = kind == 0 ? h_host_ : kind == 1 ? u_host_ : v_host_;
    auto get = [&](int i, int j) {
      i = ((i % p_.nx) + p_.nx) % p_.nx;
      j = j < 0 ? 0 : (j > rows - 1 ? rows - 1 : j);
      return f[static_cast<size_t>(j) * p_.nx + i];
    };
    return (1 - ax) * ((1 - ay) * get(i0, j0) + ay * get(i0, j0 + 1)) + ax * ((1 - ay) * get(i0 + 1, j0) + ay * get(i0 + 1, j0 + 1));
  }
};

}  // namespace lmw
