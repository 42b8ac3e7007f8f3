// The reference's native regression checks (reference tests/core.cpp:17-121) restated against the GPU library
// through the C++ façade lmw.hpp: FFT against a naive DFT, Gauss-Legendre exactness, spectral round trip,
// solid-body rotation, rest state and mass, the plane's linear gravity wave and its zonal harmonic, and the parameter
// refusals.
// Same tolerances as the reference.  Built by __graft_entry__.build() into tests/bin/core_gpu; run by
// tests/test_gpu_cpp.py.
#include <cmath>
#include <cstdio>

#include "lmw.hpp"

static int failures = 0;
static void expect(bool ok, const char *what, int line) {
  if (!ok) {
    std::fprintf(stderr, "FAIL line %d: %s\n", line, what);
    failures++;
  }
}
#define EXPECT(c) expect((c), #c, __LINE__)
#define CLOSE(a, b, tol) expect(std::fabs((a) - (b)) <= (tol), #a " ~ " #b, __LINE__)

int main() {
  const double pi = 3.14159265358979323846;
  {  // FFT vs the naive transform, and back
    const int n = 64;
    std::vector<lmw::cplx> a(n), ref(n);
    for (int k = 0; k < n; k++) a[k] = lmw::cplx(std::sin(0.3 * k) + 0.1 * k, std::cos(1.7 * k));
    for (int m = 0; m < n; m++) {
      lmw::cplx s = 0;
      for (int k = 0; k < n; k++) s += a[k] * std::exp(lmw::cplx(0, -2 * pi * m * k / n));
      ref[m] = s;
    }
    std::vector<lmw::cplx> b = a;
    lmw::fft(b, -1);
    double e = 0;
    for (int m = 0; m < n; m++) e = std::fmax(e, std::abs(b[m] - ref[m]));
    CLOSE(e, 0.0, 1e-10);
    lmw::fft(b, +1);
    e = 0;
    for (int k = 0; k < n; k++) e = std::fmax(e, std::abs(b[k] / static_cast<double>(n) - a[k]));
    CLOSE(e, 0.0, 1e-12);
    bool threw = false;
    try {
      std::vector<lmw::cplx> c(12);
      lmw::fft(c, -1);
    } catch (const std::invalid_argument &) { threw = true; }
    EXPECT(threw);
  }
  {  // Gauss-Legendre: sum of weights, exactness, ascending symmetric nodes
    std::vector<double> x, w;
    lmw::gauss_legendre(32, x, w);
    double s = 0, p10 = 0, p62 = 0;
    for (int i = 0; i < 32; i++) {
      s += w[i];
      p10 += w[i] * std::pow(x[i], 10);
      p62 += w[i] * std::pow(x[i], 62);
      if (i) EXPECT(x[i] > x[i - 1]);
      CLOSE(x[i], -x[31 - i], 1e-14);
    }
    CLOSE(s, 2.0, 1e-13);
    CLOSE(p10, 2.0 / 11, 1e-13);
    CLOSE(p62, 2.0 / 63, 1e-12);
  }
  {  // spectral round trip at T21 and solid-body rotation
    lmw::SphereParams p;
    lmw::Sphere s(21, p);
    std::vector<lmw::cplx> spec(22 * 22, 0), back;
    unsigned seed = 5;
    auto rnd = [&] {
      seed = seed * 1664525u + 1013904223u;
      return (seed >> 8) / 16777216.0 - 0.5;
    };
    for (int m = 0; m <= 21; m++)
      for (int n = m; n <= 21; n++) spec[m * 22 + n] = lmw::cplx(rnd(), m == 0 ? 0.0 : rnd());
    std::vector<double> grid;
    s.synthesize_scalar(spec, grid);
    s.analyze_scalar(grid, back);
    double e = 0;
    for (size_t i = 0; i < spec.size(); i++) e = std::fmax(e, std::abs(back[i] - spec[i]));
    CLOSE(e, 0.0, 1e-12);
    const int nlat = s.nlat(), nlon = s.nlon();
    std::vector<double> h(nlat * nlon, 1000.0), u(h), v(nlat * nlon, 0.0);
    for (int j = 0; j < nlat; j++)
      for (int k = 0; k < nlon; k++) u[j * nlon + k] = 10 * std::cos(s.lat(j));
    s.set_state_grid(h, u, v);
    CLOSE(s.vorticity_at(0.7, 1.0), 2 * 10 * std::sin(0.7) / p.radius, 1e-12);
    CLOSE(s.wind_at(0.7, 2.0, true), 10 * std::cos(0.7), 1e-9);
    CLOSE(s.wind_at(0.7, 2.0, false), 0.0, 1e-9);
    CLOSE(s.height_at(-0.3, 4.0), 1000.0, 1e-9);
    CLOSE(s.mean_height(), 1000.0, 1e-9);
    CLOSE(s.energy(), 0.5 * 1000 * (100.0 * 2.0 / 3.0) + 0.5 * p.gravity * 1e6, 1e-6 * s.energy());
  }
  {  // rest stays at rest; the step conserves mass; bad max_wave
    lmw::SphereParams p;
    lmw::Sphere s(10, p);
    std::vector<double> h(s.nlat() * s.nlon(), 5000.0), z(h.size(), 0.0);
    s.set_state_grid(h, z, z);
    s.run(86400.0, 1800.0);
    CLOSE(s.max_wind(), 0.0, 1e-9);
    CLOSE(s.height_at(0.4, 1.0), 5000.0, 1e-8);
    s.perturb(20.0, 3, 5);
    const double m0 = s.mean_height();
    s.run(43200.0, 1800.0);
    CLOSE(s.mean_height(), m0, 1e-9);
    bool threw = false;
    try {
      s.perturb(1.0, 1, 0);
    } catch (const std::invalid_argument &) { threw = true; }
    EXPECT(threw);
  }
  {  // expression operators: solid-body rotation's kinetic energy per unit depth u0^2/3 (tests/sphere_spectrum.lms:13-15)
    lmw::Sphere s(21, lmw::SphereParams{});
    s.fields("2000.0", "15.0 * cos(lat)", "0");
    CLOSE(s.kinetic_energy() / 2000.0, 75.0, 1e-7);
    bool threw = false;
    try {
      s.fields("2 +* 3", "0", "0");
    } catch (const std::runtime_error &e) { threw = std::string(e.what()) == "expression \"2 +* 3\": unexpected '*' at column 4"; }
    EXPECT(threw);
  }
  {  // plane: linear standing gravity wave after half a period, mass, bad parameters
    lmw::PlaneParams q;
    q.nx = 100;
    q.ny = 8;
    q.lx = 1.0e6;
    q.ly = 2.0e4;
    q.depth = 100;
    q.linear = true;
    lmw::Plane pl(q);
    const double k = 2 * pi / q.lx, c = std::sqrt(q.gravity * q.depth);
    CLOSE(pl.wave_speed(), c, 0.0);
    bool no_f0 = false;
    try {
      (void)pl.rossby_radius();
    } catch (const std::invalid_argument &) { no_f0 = true; }
    EXPECT(no_f0);
    for (int j = 0; j < q.ny; j++)
      for (int i = 0; i < q.nx; i++) pl.h(i, j) = q.depth + std::cos(k * pl.x_center(i));
    pl.push();
    CLOSE(pl.zonal_amplitude(1, 3), 1.0, 1e-12);
    const double x0 = pl.zonal_crest(1, 3);  // 0, or just below lx
    CLOSE(std::fmin(x0, q.lx - x0), 0.0, 1e-6);
    CLOSE(pl.harmonic_energy(1), 0.0, 0.0);
    CLOSE(pl.zeta_at_corner(5, 3), 0.0, 0.0);
    pl.run(pi / (c * k), 0);
    CLOSE(pl.h_at(0.5 * pl.dx(), 1e4) - q.depth, -std::cos(k * 0.5 * pl.dx()), 2e-3);
    CLOSE(pl.mean_depth(), q.depth, 1e-10);
    // half a period later the crest sits at lx / 2 with the same amplitude (tests/plane_gravity_wave.lms:15)
    CLOSE(pl.zonal_amplitude(1, 3), 1.0, 5e-4);
    CLOSE(pl.zonal_crest(1, 3), q.lx / 2, 2500.0);
    bool threw = false;
    try {
      lmw::PlaneParams bad;
      bad.nx = 2;
      bad.ny = 2;
      bad.lx = 1;
      bad.ly = 1;
      lmw::Plane b(bad);
    } catch (const std::invalid_argument &) { threw = true; }
    EXPECT(threw);
    threw = false;
    try {
      lmw::Sphere s(2, lmw::SphereParams{});
    } catch (const std::invalid_argument &) { threw = true; }
    EXPECT(threw);
  }
  if (failures) {
    std::fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  std::puts("core tests passed");
  return 0;
}
