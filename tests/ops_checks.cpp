// The operator table include/lmw_ops.hpp (the wx.* surface of the reference host, cpp/main.cpp:62-80,108-495).
//
//   ops_checks --schema       the schema document the table generates (compared with the reference's by sha256)
//   ops_checks --signatures   the canonical signature strings, one per line
//   ops_checks DIR            calls every Sphere and Plane operator once through the table on the GPU (files go to
//                             DIR), checks results against the façade called directly or closed forms, and the
//                             registration wrapper's refusals; prints "operator checks passed" at the end
//
// Built by __graft_entry__.build() into tests/bin/ops_checks; run by tests/test_ops_schema.py (CPU) and
// tests/test_gpu_ops.py (GPU).
#include <cmath>
#include <cstdio>
#include <set>
#include <string>
#include <vector>

#include "lmw_ops.hpp"

using lmw::ops::Args;
using lmw::ops::Value;

namespace {

int failures = 0;
std::set<std::string> called;

void fail(const std::string &what) {
  std::fprintf(stderr, "FAIL %s\n", what.c_str());
  failures++;
}

Value I(int64_t x) { return Value::integer(x); }
Value F(double x) { return Value::real(x); }
Value Str(const std::string &x) { return Value::string(x); }
Value B(bool x) { return Value::boolean(x); }

struct Script {
  lmw::ops::Host &host;
  Value invoke(const Value &self, const std::string &name, Args a) {
    const std::string type = self.kind == Value::OBJ ? self.type : "Wx";
    called.insert(type + "." + name);
    return host.call(type, name, self.kind == Value::OBJ ? self.obj : host.wx(), a);
  }
  Value wx(const std::string &name, Args a) { return invoke(Value::none(), name, std::move(a)); }
  // the call must fail with exactly `text`
  void refuses(const Value &self, const std::string &name, Args a, const std::string &text) {
    try {
      invoke(self, name, std::move(a));
      fail(name + ": no error, expected \"" + text + "\"");
    } catch (const lmw::ops::Error &e) {
      if (text != e.what()) fail(name + ": \"" + std::string(e.what()) + "\", expected \"" + text + "\"");
    }
  }
  void close(const std::string &what, double got, double want, double tol) {
    if (!(std::fabs(got - want) <= tol)) fail(what + ": " + std::to_string(got) + " vs " + std::to_string(want));
  }
};

void sphere_ops(Script &sc, const std::string &dir) {
  const double deg = lmw::ops::kDeg;
  Value s = sc.wx("sphere", {I(21)});
  Value e = sc.wx("planet", {I(21), F(6.371e6), F(7.292e-5), F(9.80616)});
  // a fresh sphere holds no state: its height error against a zero reference is 0/0, refused by the wrapper
  sc.invoke(s, "mark_reference", {});
  sc.refuses(s, "error_l2", {}, "Sphere.error_l2: the result is not finite");
  // builders return the receiver
  Value r = sc.invoke(s, "williamson6", {});
  if (r.obj != s.obj || r.type != "Sphere") fail("williamson6 returns the receiver");
  lmw::Sphere direct(21, lmw::SphereParams{});
  direct.setup(LMW_SETUP_WILLIAMSON6);
  sc.close("mean_height", sc.invoke(s, "mean_height", {}).f, direct.mean_height(), 1e-12 * direct.mean_height());
  sc.close("energy", sc.invoke(s, "energy", {}).f, direct.energy(), 1e-12 * direct.energy());
  sc.close("kinetic_energy", sc.invoke(s, "kinetic_energy", {}).f, direct.kinetic_energy(), 1e-12 * direct.kinetic_energy());
  sc.close("potential_energy", sc.invoke(s, "potential_energy", {}).f, direct.potential_energy(), 1e-12 * direct.potential_energy());
  sc.close("enstrophy", sc.invoke(s, "enstrophy", {}).f, direct.potential_enstrophy(), 1e-12 * direct.potential_enstrophy());
  sc.close("max_wind", sc.invoke(s, "max_wind", {}).f, direct.max_wind(), 1e-12 * direct.max_wind());
  sc.close("stable_dt", sc.invoke(s, "stable_dt", {}).f, direct.stable_dt(), 1e-12 * direct.stable_dt());
  sc.close("height_at", sc.invoke(s, "height_at", {F(45.0), F(30.0)}).f, direct.height_at(45 * deg, 30 * deg), 1e-9);
  sc.close("u_at", sc.invoke(s, "u_at", {F(10.0), F(200.0)}).f, direct.wind_at(10 * deg, 200 * deg, true), 1e-9);
  sc.close("v_at", sc.invoke(s, "v_at", {F(-20.0), F(95.0)}).f, direct.wind_at(-20 * deg, 95 * deg, false), 1e-9);
  sc.close("vorticity_at", sc.invoke(s, "vorticity_at", {F(60.0), F(0.0)}).f, direct.vorticity_at(60 * deg, 0.0), 1e-15);
  sc.close("wave_crest", sc.invoke(s, "wave_crest", {I(4), F(45.0)}).f, direct.wave_crest(4, 45 * deg) / deg, 1e-9);
  sc.close("mode_coefficient", sc.invoke(s, "mode_coefficient", {I(5), I(4), Str("im")}).f, direct.mode_coefficient(5, 4, true), 1e-9);
  if (sc.invoke(s, "truncation", {}).i != 21 || sc.invoke(s, "nlat", {}).i != 32 || sc.invoke(s, "nlon", {}).i != 64) fail("dimensions");
  sc.invoke(s, "run_dt", {F(7200.0), F(1200.0)});
  sc.close("time", sc.invoke(s, "time", {}).f, 7200.0, 0.0);
  sc.invoke(s, "run", {F(3600.0)});  // automatic dt
  sc.close("time after run", sc.invoke(s, "time", {}).f, 10800.0, 1e-6);
  if (!(sc.invoke(s, "error_l2", {}).f > 0) || !(sc.invoke(s, "error_linf", {}).f > 0)) fail("error norms after 3 hours");
  sc.invoke(s, "save_csv", {Str(dir + "/sphere.csv")});
  sc.invoke(s, "save_spectrum", {Str(dir + "/spectrum.txt")});
  sc.invoke(s, "save_zonal_mean", {Str(dir + "/zonal.txt")});
  sc.invoke(s, "tilt", {F(0.3)});
  sc.invoke(s, "diffusion", {F(1e14)});
  sc.invoke(s, "perturb", {F(2.0), I(7), I(10)});
  sc.invoke(s, "mode", {I(3), I(2), F(5.0), F(1000.0)});
  sc.close("mode sets the mean depth", sc.invoke(s, "mean_height", {}).f, 1000.0, 1e-9);
  sc.close("mode coefficient", sc.invoke(s, "mode_coefficient", {I(3), I(2), Str("re")}).f, 2.5, 1e-12);
  sc.invoke(e, "fields", {Str("8000.0"), Str("20.0 * cos(lat)"), Str("0")});
  sc.close("solid-body kinetic energy u0^2/3 per unit depth", sc.invoke(e, "kinetic_energy", {}).f / 8000.0, 400.0 / 3.0, 1e-7 * 400.0);
  sc.invoke(e, "orography", {Str("100.0 * cos(lat)")});
  sc.invoke(e, "williamson2", {F(0.0)});
  sc.invoke(e, "williamson5", {});
  sc.invoke(e, "tracer", {Str("1.0 + sin(lat)")});
  sc.close("tracer_at", sc.invoke(e, "tracer_at", {F(30.0), F(0.0)}).f, 1.5, 1e-9);
  sc.invoke(e, "advect_only", {});
  sc.invoke(e, "williamson1", {F(0.05)});
  sc.close("cosine bell peak", sc.invoke(e, "tracer_max", {}).f, 950.0, 150.0);
  if (!(sc.invoke(e, "tracer_min", {}).f < 0)) fail("the truncated bell undershoots");
  if (!(sc.invoke(e, "tracer_mean", {}).f > 0) || !(sc.invoke(e, "tracer_mass", {}).f > 0)) fail("tracer mean and mass");
  sc.invoke(e, "run_dt", {F(86400.0), F(3600.0)});
  if (!(sc.invoke(e, "tracer_error_l2", {}).f > 0) || !(sc.invoke(e, "tracer_error_linf", {}).f > 0)) fail("tracer errors after a day");
  // the reference's argument checks and texts
  sc.refuses(s, "tilt", {F(NAN)}, "tilt angle must be a finite number");
  sc.refuses(s, "diffusion", {F(-1.0)}, "Sphere.diffusion: coefficient must be finite and not negative");
  sc.refuses(s, "run_dt", {F(100.0), F(0.0)}, "Sphere.run_dt: dt must be positive");
  sc.refuses(s, "mode", {I(3), I(2), F(1.0), F(-5.0)}, "Sphere.mode: mean depth must be positive");
  sc.refuses(s, "mode", {I(3000000000LL), I(2), F(1.0), F(5.0)}, "Sphere.mode: integer argument out of range");
  sc.refuses(s, "perturb", {F(1.0), I(-1), I(5)}, "Sphere.perturb: seed must not be negative");
  sc.refuses(s, "height_at", {F(91.0), F(0.0)}, "Sphere.height_at: latitude must be within +-90 degrees");
  sc.refuses(s, "u_at", {F(90.0), F(0.0)}, "Sphere.u_at: latitude must be inside (-90, 90) degrees");
  sc.refuses(s, "mode_coefficient", {I(3), I(2), Str("x")}, "Sphere.mode_coefficient: part must be re or im");
  sc.refuses(s, "williamson2", {F(INFINITY)}, "angle must be a finite number");
  sc.refuses(s, "fields", {Str("2 +* 3"), Str("0"), Str("0")}, "expression \"2 +* 3\": unexpected '*' at column 4");
  sc.refuses(s, "energy", {F(1.0)}, "Sphere.energy takes 0 arguments");
  sc.refuses(s, "height_at", {I(1), F(0.0)}, "Sphere.height_at argument 1 must be float");
  sc.refuses(s, "save_csv", {Str("/nonexistent-dir/x.csv")}, "Sphere.save_csv: cannot write /nonexistent-dir/x.csv");
  sc.refuses(Value::none(), "sphere", {I(3)}, "Wx.sphere: truncation must be between 4 and 170");
  sc.refuses(Value::none(), "planet", {I(21), F(-1.0), F(0.0), F(9.8)}, "Wx.planet: radius and gravity must be positive and omega not negative");
}

void plane_ops(Script &sc, const std::string &dir) {
  Value p = sc.wx("plane", {I(64), I(32), F(6.4e5), F(3.2e5)});
  sc.invoke(p, "gravity", {F(9.81)});
  sc.invoke(p, "depth", {F(1000.0)});
  sc.invoke(p, "coriolis", {F(1e-4), F(1.6e-11)});
  sc.invoke(p, "linear", {});
  sc.invoke(p, "nonlinear", {});
  sc.invoke(p, "viscosity", {F(10.0)});
  sc.invoke(p, "drag", {F(1e-7)});
  sc.close("wave_speed", sc.invoke(p, "wave_speed", {}).f, std::sqrt(9.81 * 1000.0), 1e-12);
  sc.close("rossby_radius", sc.invoke(p, "rossby_radius", {}).f, std::sqrt(9.81 * 1000.0) / 1e-4, 1e-6);
  sc.close("mean_depth at rest", sc.invoke(p, "mean_depth", {}).f, 1000.0, 1e-10);
  sc.invoke(p, "fields", {Str("cos(2 * 3.141592653589793 * x / lx)"), Str("0.1"), Str("0")});
  sc.close("h_at after fields", sc.invoke(p, "h_at", {F(5000.0), F(1.6e5)}).f, 1000.0 + std::cos(2 * lmw::ops::kPi * 5000.0 / 6.4e5), 1e-9);
  sc.close("u_at after fields", sc.invoke(p, "u_at", {F(1.0e5), F(1.6e5)}).f, 0.1, 1e-12);
  sc.close("zonal_amplitude", sc.invoke(p, "zonal_amplitude", {I(1), I(5)}).f, 1.0, 1e-9);
  sc.refuses(p, "depth", {F(500.0)}, "Plane.depth: set the depth before fields() or perturb()");
  sc.invoke(p, "perturb", {F(0.01), I(3)});
  const double dt = sc.invoke(p, "stable_dt", {}).f;
  sc.invoke(p, "run_dt", {F(10 * 0.5 * dt), F(0.5 * dt)});
  sc.invoke(p, "run", {F(600.0)});
  if (!(sc.invoke(p, "time", {}).f > 0)) fail("plane time");
  if (!(sc.invoke(p, "energy", {}).f > 0) || !(sc.invoke(p, "max_speed", {}).f > 0)) fail("plane energy and speed");
  (void)sc.invoke(p, "pv_anomaly_rms", {});
  (void)sc.invoke(p, "v_at", {F(3.0e5), F(1.0e5)});
  (void)sc.invoke(p, "zonal_crest", {I(1), I(5)});
  (void)sc.invoke(p, "v_zonal_crest", {I(1), I(5)});
  (void)sc.invoke(p, "harmonic_energy", {I(1)});
  sc.invoke(p, "save_csv", {Str(dir + "/plane.csv")});
  sc.invoke(p, "save_row", {I(3), Str(dir + "/row.txt")});
  sc.invoke(p, "save_column", {I(5), Str(dir + "/column.txt")});
  sc.refuses(p, "save_row", {I(32), Str(dir + "/row.txt")}, "Plane.save_row: row outside the grid");
  sc.refuses(p, "save_column", {I(-1), Str(dir + "/c.txt")}, "Plane.save_column: column outside the grid");
  sc.refuses(p, "h_at", {F(-1.0), F(0.0)}, "Plane.h_at: point outside the domain");
  sc.refuses(p, "gravity", {F(0.0)}, "Plane.gravity: must be positive");
  sc.refuses(p, "viscosity", {F(NAN)}, "Plane.viscosity: must be finite and not negative");
  sc.refuses(p, "coriolis", {F(1.0), F(INFINITY)}, "beta must be a finite number");
  sc.refuses(p, "perturb", {F(1.0), I(-3)}, "Plane.perturb: seed must not be negative");
  sc.refuses(p, "fields", {Str("1/0"), Str("0"), Str("0")}, "Plane.fields: the height expression is not finite");
  sc.refuses(Value::none(), "plane", {I(2), I(8), F(1.0), F(1.0)}, "Wx.plane: need 4..4000 cells each way and at most 8 million cells");
  Value q = sc.wx("plane", {I(16), I(8), F(1.6e5), F(8e4)});
  sc.refuses(q, "rossby_radius", {}, "Plane.rossby_radius: f0 is zero");
}

void wx_ops(Script &sc, const std::string &dir) {
  sc.wx("assert_close", {F(1.0), F(1.05), F(0.1)});
  sc.refuses(Value::none(), "assert_close", {F(1.0), F(2.0), F(0.5)}, "assert_close failed: 1 vs 2 (tolerance 0.5)");
  sc.wx("assert_true", {B(true)});
  sc.refuses(Value::none(), "assert_true", {B(false)}, "assert_true failed");
  sc.wx("print", {Str("operator table: print")});
  sc.wx("save_text", {Str(dir + "/text.txt"), Str("hello\n")});
  if (sc.wx("version", {}).s != "lmweather 0.1.4") fail("version");
}

}  // namespace

int main(int argc, char **argv) {
  lmw::ops::Host host;
  const std::string arg = argc > 1 ? argv[1] : "";
  if (arg == "--schema") {
    std::fputs(host.schema_json().c_str(), stdout);
    return 0;
  }
  if (arg == "--signatures") {
    for (const auto &s : host.signatures()) std::printf("%s\n", s.c_str());
    return 0;
  }
  const std::string dir = arg.empty() ? "/tmp" : arg;
  Script sc{host};
  try {
    wx_ops(sc, dir);
    sphere_ops(sc, dir);
    plane_ops(sc, dir);
  } catch (const std::exception &e) {
    fail(std::string("unexpected error: ") + e.what());
  }
  for (const auto &op : host.operators()) {
    const std::string key = op.type + "." + op.name;
    if (!called.count(key)) fail("operator never called: " + key);
  }
  if (failures) {
    std::fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  std::printf("operator checks passed (%zu operators)\n", host.operators().size());
  return 0;
}
