// Checks of the C++ facade include/lmw.hpp, the class API a native caller of the reference's Sphere and Plane
// (reference cpp/sphere.hpp:27-121, cpp/plane.hpp:20-68) links against, run on the GPU library.
//
// Each check states a closed form or an exact property and its bound.  The bounds are the ones the reference's own
// native test asserts for the same property (reference tests/core.cpp:17-121: 1e-10 for a transform against its
// exact value, 1e-12/1e-13 for quadrature and round trips, 1e-9 for winds and heights, 1e-6 relative for energy,
// 2e-3 for the gravity wave after half a period); the inputs are chosen here.  Built by __graft_entry__.build() into
// tests/bin/facade_checks and run by tests/test_gpu_cpp.py.
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "lmw.hpp"

namespace {

const double kPi = 3.14159265358979323846;

// One named check: a measured deviation against its bound, or a condition.
struct Ledger {
  int checks = 0, failed = 0;
  std::string section;
  void within(const char *what, double deviation, double bound) {
    checks++;
    if (!(std::fabs(deviation) <= bound)) {
      failed++;
      std::fprintf(stderr, "FAIL [%s] %s: |deviation| %.3e > %.3e\n", section.c_str(), what, std::fabs(deviation), bound);
    }
  }
  void holds(const char *what, bool ok) {
    checks++;
    if (!ok) {
      failed++;
      std::fprintf(stderr, "FAIL [%s] %s\n", section.c_str(), what);
    }
  }
  // the call throws E whose what() equals `text` (empty: any text)
  template <class E>
  void refuses(const char *what, const std::function<void()> &call, const std::string &text = "") {
    checks++;
    try {
      call();
    } catch (const E &e) {
      if (text.empty() || text == e.what()) return;
      failed++;
      std::fprintf(stderr, "FAIL [%s] %s: message \"%s\", expected \"%s\"\n", section.c_str(), what, e.what(), text.c_str());
      return;
    } catch (const std::exception &e) {
      failed++;
      std::fprintf(stderr, "FAIL [%s] %s: wrong exception type (\"%s\")\n", section.c_str(), what, e.what());
      return;
    }
    failed++;
    std::fprintf(stderr, "FAIL [%s] %s: no exception\n", section.c_str(), what);
  }
};

// A sum of three complex tones has an exactly known discrete spectrum: n A_j at bin f_j, zero elsewhere.
void fft_of_tones(Ledger &L) {
  const int n = 256;
  const int bins[3] = {3, 17, n - 5};
  const lmw::cplx amp[3] = {lmw::cplx(1.5, -0.5), lmw::cplx(-0.25, 2.0), lmw::cplx(0.75, 0.0)};
  std::vector<lmw::cplx> x(n, 0.0);
  for (int k = 0; k < n; k++)
    for (int t = 0; t < 3; t++) x[k] += amp[t] * std::polar(1.0, 2 * kPi * bins[t] * k / n);
  std::vector<lmw::cplx> X = x;
  lmw::fft(X, -1);
  double worst = 0;
  for (int f = 0; f < n; f++) {
    lmw::cplx want = 0.0;
    for (int t = 0; t < 3; t++)
      if (bins[t] == f) want = amp[t] * static_cast<double>(n);
    worst = std::fmax(worst, std::abs(X[f] - want));
  }
  L.within("forward transform of three tones", worst, 1e-10);
  lmw::fft(X, +1);  // unnormalized: n x
  worst = 0;
  for (int k = 0; k < n; k++) worst = std::fmax(worst, std::abs(X[k] / static_cast<double>(n) - x[k]));
  L.within("inverse recovers the signal", worst, 1e-12);
  L.refuses<std::invalid_argument>("length 96", [] {
    std::vector<lmw::cplx> y(96);
    lmw::fft(y, -1);
  }, "fft length must be a power of two");
}

// An n-point Gauss rule integrates every polynomial of degree <= 2n - 1 exactly: the Legendre polynomials P_1..P_79
// (Bonnet's recurrence at the nodes) to 0 and x^78 to 2/79.
void gauss_rule(Ledger &L) {
  const int n = 40;
  std::vector<double> x, w;
  lmw::gauss_legendre(n, x, w);
  double sw = 0, worst_p = 0, x78 = 0;
  std::vector<double> I(2 * n, 0.0);
  for (int i = 0; i < n; i++) {
    sw += w[i];
    x78 += w[i] * std::pow(x[i], 78);
    double pm = 1.0, p = x[i];  // P_0, P_1
    I[1] += w[i] * p;
    for (int k = 2; k < 2 * n; k++) {
      const double pn = ((2.0 * k - 1) * x[i] * p - (k - 1.0) * pm) / k;
      pm = p;
      p = pn;
      I[k] += w[i] * p;
    }
  }
  for (int k = 1; k < 2 * n; k++) worst_p = std::fmax(worst_p, std::fabs(I[k]));
  L.within("weights sum to 2", sw - 2.0, 1e-13);
  L.within("P_1..P_79 integrate to 0", worst_p, 1e-13);
  L.within("x^78 integrates to 2/79", x78 - 2.0 / 79.0, 1e-12);
  bool ordered = true;
  for (int i = 1; i < n; i++) ordered = ordered && x[i] > x[i - 1] && std::fabs(x[i] + x[n - 1 - i]) <= 1e-14;
  L.holds("nodes ascending and symmetric", ordered);
}

// analyze(synthesize(c)) = c for any band-limited spectrum.
void round_trip(Ledger &L) {
  const int N = 31;
  lmw::Sphere s(N, lmw::SphereParams{});
  std::vector<lmw::cplx> c((N + 1) * (N + 1), 0.0), back;
  for (int m = 0; m <= N; m++)
    for (int n = m; n <= N; n++) c[m * (N + 1) + n] = lmw::cplx(std::cos(0.37 * m + 1.3 * n), m ? std::sin(0.11 * m * n + 0.2) : 0.0);
  std::vector<double> g;
  s.synthesize_scalar(c, g);
  s.analyze_scalar(g, back);
  double worst = 0;
  for (size_t i = 0; i < c.size(); i++) worst = std::fmax(worst, std::abs(back[i] - c[i]));
  L.within("T31 synthesis then analysis", worst, 1e-12);
}

// Solid-body rotation u = u0 cos(lat) on a layer of depth H: zeta = 2 u0 sin(lat) / a, v = 0, and the area mean of
// the energy h |v|^2 / 2 + g h^2 / 2 is H u0^2 / 3 + g H^2 / 2 (the area mean of cos^2 is 2/3).
void solid_body(Ledger &L) {
  const double u0 = 25.0, H = 3000.0;
  lmw::SphereParams p;
  lmw::Sphere s(30, p);
  const int J = s.nlat(), K = s.nlon();
  std::vector<double> h(J * K, H), u(J * K), v(J * K, 0.0);
  for (int j = 0; j < J; j++)
    for (int k = 0; k < K; k++) u[j * K + k] = u0 * std::cos(s.lat(j));
  s.set_state_grid(h, u, v);
  const double pts[3][2] = {{0.9, 0.3}, {-0.45, 2.5}, {0.05, 5.9}};
  for (const auto &q : pts) {
    L.within("vorticity 2 u0 sin(lat) / a", s.vorticity_at(q[0], q[1]) - 2 * u0 * std::sin(q[0]) / p.radius, 1e-12);
    L.within("zonal wind u0 cos(lat)", s.wind_at(q[0], q[1], true) - u0 * std::cos(q[0]), 1e-9);
    L.within("meridional wind 0", s.wind_at(q[0], q[1], false), 1e-9);
    L.within("height H", s.height_at(q[0], q[1]) - H, 1e-9);
  }
  L.within("mean height H", s.mean_height() - H, 1e-9);
  const double e = H * u0 * u0 / 3.0 + 0.5 * p.gravity * H * H;
  L.within("energy H u0^2/3 + g H^2/2", s.energy() - e, 1e-6 * e);
}

// A resting layer stays at rest, and stepping conserves the mean height of a perturbed layer.
void rest_and_mass(Ledger &L) {
  lmw::Sphere s(12, lmw::SphereParams{});
  std::vector<double> h(s.nlat() * s.nlon(), 4000.0), zero(h.size(), 0.0);
  s.set_state_grid(h, zero, zero);
  s.run(86400.0, 1200.0);
  L.within("no wind after a day at rest", s.max_wind(), 1e-9);
  L.within("height unchanged", s.height_at(-0.6, 3.3) - 4000.0, 1e-8);
  s.perturb(15.0, 7, 6);
  const double m0 = s.mean_height();
  s.run(43200.0, 1200.0);
  L.within("mean height conserved by the step", s.mean_height() - m0, 1e-9);
  L.refuses<std::invalid_argument>("max_wave 0", [&] { s.perturb(1.0, 1, 0); }, "max_wave must be between 1 and the truncation");
}

// The expression operator: solid-body rotation of 20 m/s carries u0^2/3 of kinetic energy per unit depth; strings
// that do not compile are refused with the compiler's text and column.
void expressions(Ledger &L) {
  lmw::Sphere s(21, lmw::SphereParams{});
  s.fields("2500.0", "20.0 * cos(lat)", "0");
  L.within("kinetic energy per unit depth u0^2/3", s.kinetic_energy() / 2500.0 - 400.0 / 3.0, 1e-7 * 400.0 / 3.0);
  L.refuses<std::runtime_error>("unclosed parenthesis", [&] { s.fields("(1 + lat", "0", "0"); },
                                "expression \"(1 + lat\": expected ')' at column 9");
  L.refuses<std::runtime_error>("unknown character", [&] { s.fields("3 $ 2", "0", "0"); },
                                "expression \"3 $ 2\": unexpected '$' at column 3");
}

// A linear standing gravity wave cos(k x) of zonal wavenumber 2 on a channel: after half a period the height is
// -cos(k x); its crest, fixed by the symmetry of a standing wave, moves from x = 0 to x = lx / 4.
void plane_wave(Ledger &L) {
  lmw::PlaneParams q;
  q.nx = 128;
  q.ny = 6;
  q.lx = 2.0e6;
  q.ly = 3.0e4;
  q.depth = 50.0;
  q.linear = true;
  lmw::Plane pl(q);
  const double k = 2 * 2 * kPi / q.lx, c = std::sqrt(q.gravity * q.depth);
  L.within("wave speed sqrt(g H)", pl.wave_speed() - c, 0.0);
  L.refuses<std::invalid_argument>("rossby radius without rotation", [&] { (void)pl.rossby_radius(); });
  for (int j = 0; j < q.ny; j++)
    for (int i = 0; i < q.nx; i++) pl.h(i, j) = q.depth + std::cos(k * pl.x_center(i));
  pl.push();
  L.within("harmonic 2 has amplitude 1", pl.zonal_amplitude(2, 2) - 1.0, 1e-12);
  L.within("no kinetic energy at rest", pl.harmonic_energy(2), 0.0);
  L.within("no corner vorticity at rest", pl.zeta_at_corner(7, 2), 0.0);
  pl.run(kPi / (c * k), 0);
  L.within("h - H = -cos(k x) after half a period", pl.h_at(3.5 * pl.dx(), 1.5e4) - q.depth + std::cos(k * 3.5 * pl.dx()), 2e-3);
  L.within("mean depth conserved", pl.mean_depth() - q.depth, 1e-10);
  L.within("amplitude kept", pl.zonal_amplitude(2, 2) - 1.0, 5e-4);
  L.within("crest at lx / 4", pl.zonal_crest(2, 2) - q.lx / 4, 1e-6 * q.lx);
}

void refusals(Ledger &L) {
  L.refuses<std::invalid_argument>("plane 3 x 3", [] {
    lmw::PlaneParams bad;
    bad.nx = 3;
    bad.ny = 3;
    bad.lx = 1;
    bad.ly = 1;
    lmw::Plane b(bad);
  });
  L.refuses<std::invalid_argument>("truncation 3", [] { lmw::Sphere s(3, lmw::SphereParams{}); },
                                   "truncation must be between 4 and 170");
  L.refuses<std::invalid_argument>("truncation 171", [] { lmw::Sphere s(171, lmw::SphereParams{}); },
                                   "truncation must be between 4 and 170");
}

}  // namespace

int main() {
  const std::pair<const char *, void (*)(Ledger &)> sections[] = {
      {"fft", fft_of_tones},       {"gauss", gauss_rule},           {"round trip", round_trip},
      {"solid body", solid_body},  {"rest", rest_and_mass},         {"expressions", expressions},
      {"plane wave", plane_wave},  {"refusals", refusals},
  };
  Ledger L;
  for (const auto &s : sections) {
    L.section = s.first;
    try {
      s.second(L);
    } catch (const std::exception &e) {
      L.failed++;
      std::fprintf(stderr, "FAIL [%s] unexpected exception: %s\n", s.first, e.what());
    }
  }
  if (L.failed) {
    std::fprintf(stderr, "%d of %d checks failed\n", L.failed, L.checks);
    return 1;
  }
  std::printf("facade checks passed (%d checks)\n", L.checks);
  return 0;
}
