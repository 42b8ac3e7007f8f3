// Text output of the sphere and the plane: the save_* operators of the reference's registration layer
// (cpp/main.cpp:83-92 `g` and `save`, :328-335 Sphere.save_spectrum, :357-371 Sphere.save_csv, :373-383
// Sphere.save_zonal_mean, :464-493 Plane.save_csv / save_row / save_column), written over this library's own C ABI.
// Every number is printed with "%.10g" and the layouts are the reference's: one line per point, and for the grids a
// blank line after each latitude or plane row (the gnuplot pm3d layout).  The fields come back through the ABI's
// getters; the formatting is host work.  A file's bytes are a function of the downloaded doubles alone, so reruns
// of a deterministic run write identical files.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "common.h"

namespace {
using lmw::invalid;
using lmw::runtime;

constexpr double kPi = 3.14159265358979323846;
constexpr double kDeg = kPi / 180.0;

// a status from a call of this ABI back into an exception, keeping its code and message
void check(lmw_status s) {
  if (s != LMW_OK) throw lmw::Error(s, lmw_last_error());
}

// "%.10g" (cpp/main.cpp:84)
void put(std::string &t, double x) {
  char b[48];
  const int n = std::snprintf(b, sizeof b, "%.10g", x);
  t.append(b, static_cast<size_t>(n));
}
void put_line(std::string &t, std::initializer_list<double> xs) {
  bool first = true;
  for (double x : xs) {
    if (!first) t += ' ';
    put(t, x);
    first = false;
  }
  t += '\n';
}

// cpp/main.cpp:86-92: the whole text, or the error "<who>: cannot write <path>"
void save(const char *path, const std::string &text, const char *who) {
  if (!path) invalid(std::string(who) + ": null path");
  std::FILE *f = std::fopen(path, "wb");
  if (!f) runtime(std::string(who) + ": cannot write " + path);
  const size_t w = std::fwrite(text.data(), 1, text.size(), f);
  const int flushed = std::fflush(f);
  const int closed = std::fclose(f);
  if (w != text.size() || flushed != 0 || closed != 0) runtime(std::string(who) + ": cannot write " + path);
}

struct SphereGrids {
  int N = 0, J = 0, K = 0;
  std::vector<double> mu, w;
  size_t n() const { return static_cast<size_t>(J) * K; }
};
SphereGrids sphere_dims(lmw_sphere *s) {
  SphereGrids d;
  check(lmw_sphere_dims(s, &d.N, &d.J, &d.K));
  d.mu.resize(d.J);
  d.w.resize(d.J);
  check(lmw_sphere_gauss(s, d.mu.data(), d.w.data()));
  return d;
}
std::vector<double> sphere_grid(lmw_sphere *s, const SphereGrids &d, int field) {
  std::vector<double> g(d.n());
  check(lmw_sphere_get_grid(s, field, g.data(), g.size()));
  return g;
}
double lat_deg(const SphereGrids &d, int j) { return std::asin(d.mu[j]) / kDeg; }           // sphere.cpp:120
double lon_deg(const SphereGrids &d, int k) { return 2 * kPi * k / d.K / kDeg; }            // sphere.cpp:121

struct PlaneView {
  lmw_plane_params p{};
  int batch = 0, precision = 0;
  std::vector<double> h, u, v;
  double dx() const { return p.lx / p.nx; }  // plane.hpp:26-31
  double dy() const { return p.ly / p.ny; }
  double x_center(int i) const { return (i + 0.5) * dx(); }
  double y_center(int j) const { return (j + 0.5) * dy(); }
  double H(int i, int j) const { return h[static_cast<size_t>(j) * p.nx + i]; }
  double U(int i, int j) const { return u[static_cast<size_t>(j) * p.nx + i]; }
  double V(int i, int j) const { return v[static_cast<size_t>(j) * p.nx + i]; }
  // plane.cpp:220-226
  double zeta_at_corner(int i, int j) const {
    const int nx = p.nx, ny = p.ny;
    const int im = (i + nx - 1) % nx;
    const double dvdx = (V(i % nx, j) - V(im, j)) / dx();
    const double uj = j < ny ? U(i % nx, j) : U(i % nx, ny - 1), ujm = j > 0 ? U(i % nx, j - 1) : U(i % nx, 0);
    return dvdx - (uj - ujm) / dy();
  }
};
PlaneView plane_view(lmw_plane *pl, int member, bool winds) {
  PlaneView m;
  check(lmw_plane_get_params(pl, &m.p, &m.batch, &m.precision));
  if (member < 0 || member >= m.batch) invalid("member index out of range");
  const size_t nh = static_cast<size_t>(m.p.nx) * m.p.ny, nv = static_cast<size_t>(m.p.nx) * (m.p.ny + 1);
  m.h.resize(nh);
  check(lmw_plane_get_field(pl, member, LMW_PLANE_H, m.h.data(), nh));
  if (winds) {
    m.u.resize(nh);
    m.v.resize(nv);
    check(lmw_plane_get_field(pl, member, LMW_PLANE_U, m.u.data(), nh));
    check(lmw_plane_get_field(pl, member, LMW_PLANE_V, m.v.data(), nv));
  }
  return m;
}
}  // namespace

extern "C" {

lmw_status lmw_sphere_save_csv(lmw_sphere *s, const char *path) {
  return lmw::guard([&] {
    const SphereGrids d = sphere_dims(s);
    const std::vector<double> hh = sphere_grid(s, d, LMW_GRID_H), uu = sphere_grid(s, d, LMW_GRID_U), vv = sphere_grid(s, d, LMW_GRID_V),
                              zz = sphere_grid(s, d, LMW_GRID_VORTICITY), hs = sphere_grid(s, d, LMW_GRID_OROGRAPHY);
    int has_tracer = 0, advect_only = 0;
    check(lmw_sphere_tracer_flags(s, &has_tracer, &advect_only));
    const std::vector<double> qq = has_tracer ? sphere_grid(s, d, LMW_GRID_TRACER) : std::vector<double>(d.n(), 0.0);
    std::string t;
    t.reserve(d.n() * 8 * 14);
    for (int j = 0; j < d.J; j++) {
      const double lat = lat_deg(d, j);
      for (int k = 0; k < d.K; k++) {
        const size_t i = static_cast<size_t>(j) * d.K + k;
        put_line(t, {lon_deg(d, k), lat, hh[i], uu[i], vv[i], zz[i], qq[i], hs[i]});
      }
      t += '\n';
    }
    save(path, t, "Sphere.save_csv");
  });
}

lmw_status lmw_sphere_save_spectrum(lmw_sphere *s, const char *path) {
  return lmw::guard([&] {
    int N = 0, J = 0, K = 0;
    check(lmw_sphere_dims(s, &N, &J, &K));
    std::vector<double> ke(N + 1), pe(N + 1);
    check(lmw_sphere_spectrum(s, ke.data(), pe.data(), ke.size()));
    std::string t;
    for (int n = 1; n <= N; n++) {
      t += std::to_string(n);
      t += ' ';
      put_line(t, {ke[n], pe[n]});
    }
    save(path, t, "Sphere.save_spectrum");
  });
}

lmw_status lmw_sphere_save_zonal_mean(lmw_sphere *s, const char *path) {
  return lmw::guard([&] {
    const SphereGrids d = sphere_dims(s);
    const std::vector<double> hh = sphere_grid(s, d, LMW_GRID_H), uu = sphere_grid(s, d, LMW_GRID_U);
    std::string t;
    for (int j = 0; j < d.J; j++) {
      double sh = 0, su = 0;  // in longitude order, like the reference
      for (int k = 0; k < d.K; k++) {
        sh += hh[static_cast<size_t>(j) * d.K + k];
        su += uu[static_cast<size_t>(j) * d.K + k];
      }
      put_line(t, {lat_deg(d, j), sh / d.K, su / d.K});
    }
    save(path, t, "Sphere.save_zonal_mean");
  });
}

lmw_status lmw_plane_save_csv(lmw_plane *p, int member, const char *path) {
  return lmw::guard([&] {
    const PlaneView m = plane_view(p, member, true);
    const int nx = m.p.nx, ny = m.p.ny;
    std::string t;
    t.reserve(static_cast<size_t>(nx) * ny * 6 * 14);
    for (int j = 0; j < ny; j++) {
      for (int i = 0; i < nx; i++) {
        const double uc = 0.5 * (m.U(i, j) + m.U((i + 1) % nx, j)), vc = 0.5 * (m.V(i, j) + m.V(i, j + 1));
        const double z = 0.25 * (m.zeta_at_corner(i, j) + m.zeta_at_corner(i + 1, j) + m.zeta_at_corner(i, j + 1) + m.zeta_at_corner(i + 1, j + 1));
        put_line(t, {m.x_center(i), m.y_center(j), m.H(i, j), uc, vc, z});
      }
      t += '\n';
    }
    save(path, t, "Plane.save_csv");
  });
}

lmw_status lmw_plane_save_row(lmw_plane *p, int member, int row, const char *path) {
  return lmw::guard([&] {
    const PlaneView m = plane_view(p, member, false);
    if (row < 0 || row >= m.p.ny) invalid("Plane.save_row: row outside the grid");
    std::string t;
    for (int i = 0; i < m.p.nx; i++) put_line(t, {m.x_center(i), m.H(i, row)});
    save(path, t, "Plane.save_row");
  });
}

lmw_status lmw_plane_save_column(lmw_plane *p, int member, int column, const char *path) {
  return lmw::guard([&] {
    const PlaneView m = plane_view(p, member, false);
    if (column < 0 || column >= m.p.nx) invalid("Plane.save_column: column outside the grid");
    std::string t;
    for (int j = 0; j < m.p.ny; j++) put_line(t, {m.y_center(j), m.H(column, j)});
    save(path, t, "Plane.save_column");
  });
}

}  // extern "C"
