// lmw_ops.hpp -- the wx.* operator surface of lmweather's script host over the C++ façade lmw.hpp: the 81 host
// methods (Wx 8, Sphere 44, Plane 29) that the reference registers with its VM (reference cpp/main.cpp:108-495), in
// the same order, with the same parameter and return types, the same argument checks and texts, and the
// registration wrapper's behaviour (cpp/main.cpp:62-80): each argument's kind is checked, every exception becomes an
// ops::Error carrying its message, and a float result that is not finite is refused.  Header-only; a host that binds
// a script VM registers operators()[i].fn under (type, name, params, returns).
//
// schema_json() writes the schema document the reference generates from its registrations
// (scripts/gen-schema.py:8-17): types in registration order, two-space indentation, no functions, the context
// wx: Wx.  signatures() lists the canonical strings of the registry ("method Type.name(params) -> returns" and
// "context wx: Wx", sorted), the form of `--lmcode-signatures`.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "lmw.hpp"

namespace lmw {
namespace ops {

// The error a script sees (the reference's dslvm::Error): the message only.
struct Error : std::runtime_error {
  explicit Error(const std::string &m) : std::runtime_error(m) {}
};

// A script value: none, int (64-bit), float, bool, string or an object of a named type.
struct Value {
  enum Kind { NONE, INT, FLOAT, BOOL, STR, OBJ };
  Kind kind = NONE;
  int64_t i = 0;
  double f = 0;
  bool b = false;
  std::string s, type;
  void *obj = nullptr;
  static Value none() { return Value{}; }
  static Value integer(int64_t x) {
    Value v;
    v.kind = INT;
    v.i = x;
    return v;
  }
  static Value real(double x) {
    Value v;
    v.kind = FLOAT;
    v.f = x;
    return v;
  }
  static Value boolean(bool x) {
    Value v;
    v.kind = BOOL;
    v.b = x;
    return v;
  }
  static Value string(std::string x) {
    Value v;
    v.kind = STR;
    v.s = std::move(x);
    return v;
  }
  static Value object(std::string t, void *p) {
    Value v;
    v.kind = OBJ;
    v.type = std::move(t);
    v.obj = p;
    return v;
  }
};
using Args = std::vector<Value>;
using Method = std::function<Value(void *self, Args &)>;

struct Op {
  std::string type, name;
  std::vector<std::string> params;
  std::string returns;
  Method fn;
};

// ---------------------------------------------------------------- argument checks (cpp/main.cpp:82-106)
constexpr double kPi = 3.14159265358979323846;
constexpr double kDeg = kPi / 180.0;
inline void need(bool ok, const std::string &msg) {
  if (!ok) throw Error(msg);
}
inline double finite(double x, const std::string &what) {
  need(std::isfinite(x), what + " must be a finite number");
  return x;
}
// an int argument used as a C int: values a cast would wrap are refused
inline int narrow(int64_t v, const std::string &who) {
  need(v >= -1000000000 && v <= 1000000000, who + ": integer argument out of range");
  return static_cast<int>(v);
}
inline double positive(double x, const std::string &msg) {
  need(std::isfinite(x) && x > 0, msg);
  return x;
}
inline double latitude(double deg, const std::string &who) {
  need(std::fabs(deg) <= 90.0, who + ": latitude must be within +-90 degrees");
  return deg * kDeg;
}
inline std::string full_digits(double x) {  // "%.17g": the numbers of an assert_close failure
  char b[48];
  std::snprintf(b, sizeof b, "%.17g", x);
  return b;
}
inline void write_text(const std::string &path, const std::string &text, const std::string &who) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw Error(who + ": cannot write " + path);
  f << text;
  f.flush();
  if (!f) throw Error(who + ": cannot write " + path);
}

// The objects a script holds.  A plane keeps its parameters and whether fields were set (depth() resets them).
struct SphereObj {
  std::unique_ptr<Sphere> model;
};
struct PlaneObj {
  PlaneParams params;
  std::unique_ptr<Plane> model;
  bool started = false;
};
struct WxObj {};

class Host {
 public:
  static constexpr size_t kObjectLimit = 100000;  // objects per type, as the reference's arenas

  Host() { register_all(); }
  Host(const Host &) = delete;
  Host &operator=(const Host &) = delete;

  const std::vector<Op> &operators() const { return ops_; }
  void *wx() { return &wx_; }  // the context object `wx`

  const Op *find(const std::string &type, const std::string &name) const {
    for (const Op &op : ops_)
      if (op.type == type && op.name == name) return &op;
    return nullptr;
  }

  // One call through the registration wrapper: argument kinds, exception translation, finite float results.
  Value call(const Op &op, void *self, Args &a) const {
    const std::string who = op.type + "." + op.name;
    need(a.size() == op.params.size(), who + " takes " + std::to_string(op.params.size()) + " arguments");
    for (size_t k = 0; k < op.params.size(); k++) {
      const std::string &want = op.params[k];
      const Value &v = a[k];
      const bool ok = want == "float"    ? v.kind == Value::FLOAT
                      : want == "int"    ? v.kind == Value::INT
                      : want == "string" ? v.kind == Value::STR
                      : want == "bool"   ? v.kind == Value::BOOL
                                         : v.kind == Value::OBJ && v.obj && v.type == want;
      if (!ok) throw Error(who + " argument " + std::to_string(k + 1) + " must be " + want);
    }
    Value out;
    try {
      out = op.fn(self, a);
    } catch (const Error &) {
      throw;
    } catch (const std::exception &e) {
      throw Error(e.what());
    }
    if (op.returns == "float" && out.kind == Value::FLOAT && !std::isfinite(out.f)) throw Error(who + ": the result is not finite");
    return out;
  }
  Value call(const std::string &type, const std::string &name, void *self, Args &a) const {
    const Op *op = find(type, name);
    need(op != nullptr, "no operator " + type + "." + name);
    return call(*op, self, a);
  }

  std::string schema_json() const {
    // json.dumps(indent=2) of {"dsl.schema": 2, "name": ..., "types": {T: {"methods": {name: [{"params": [...],
    // "returns": R}]}}}, "functions": {}, "context": {"wx": "Wx"}} plus a final newline
    auto q = [](const std::string &x) { return "\"" + x + "\""; };
    std::vector<std::string> types;
    for (const Op &op : ops_)
      if (std::find(types.begin(), types.end(), op.type) == types.end()) types.push_back(op.type);
    std::string o = "{\n  \"dsl.schema\": 2,\n  \"name\": \"lmweather\",\n  \"types\": {\n";
    for (size_t t = 0; t < types.size(); t++) {
      o += "    " + q(types[t]) + ": {\n      \"methods\": {\n";
      std::vector<const Op *> ms;
      for (const Op &op : ops_)
        if (op.type == types[t]) ms.push_back(&op);
      for (size_t k = 0; k < ms.size(); k++) {
        o += "        " + q(ms[k]->name) + ": [\n          {\n            \"params\": ";
        if (ms[k]->params.empty()) {
          o += "[],\n";
        } else {
          o += "[\n";
          for (size_t p = 0; p < ms[k]->params.size(); p++)
            o += "              " + q(ms[k]->params[p]) + (p + 1 < ms[k]->params.size() ? ",\n" : "\n");
          o += "            ],\n";
        }
        o += "            \"returns\": " + q(ms[k]->returns) + "\n          }\n        ]" + (k + 1 < ms.size() ? ",\n" : "\n");
      }
      o += std::string("      }\n    }") + (t + 1 < types.size() ? ",\n" : "\n");
    }
    o += "  },\n  \"functions\": {},\n  \"context\": {\n    \"wx\": \"Wx\"\n  }\n}\n";
    return o;
  }

  std::vector<std::string> signatures() const {
    std::vector<std::string> out{"context wx: Wx"};
    for (const Op &op : ops_) {
      std::string s = "method " + op.type + "." + op.name + "(";
      for (size_t p = 0; p < op.params.size(); p++) s += (p ? ", " : "") + op.params[p];
      out.push_back(s + ") -> " + op.returns);
    }
    std::sort(out.begin(), out.end());
    return out;
  }

 private:
  std::vector<Op> ops_;
  WxObj wx_;
  std::vector<std::unique_ptr<SphereObj>> spheres_;
  std::vector<std::unique_ptr<PlaneObj>> planes_;

  void def(const char *type, const char *name, std::vector<std::string> params, const char *returns, Method fn) {
    ops_.push_back(Op{type, name, std::move(params), returns, std::move(fn)});
  }
  static Sphere &S(void *p) { return *static_cast<SphereObj *>(p)->model; }
  static PlaneObj &P(void *p) { return *static_cast<PlaneObj *>(p); }
  static Value self_sphere(void *p) { return Value::object("Sphere", p); }
  static Value self_plane(void *p) { return Value::object("Plane", p); }
  template <class T> Value keep(std::vector<std::unique_ptr<T>> &arena, const char *type, std::unique_ptr<T> o) {
    if (arena.size() >= kObjectLimit) throw Error(std::string(type) + " object limit exceeded");
    arena.push_back(std::move(o));
    return Value::object(type, arena.back().get());
  }
  Value new_sphere(int64_t n, const SphereParams &sp) {
    need(n >= 4 && n <= 170, "Wx.sphere: truncation must be between 4 and 170");
    need(std::isfinite(sp.radius) && std::isfinite(sp.gravity) && std::isfinite(sp.omega) && sp.radius > 0 && sp.gravity > 0 &&
             sp.omega >= 0,
         "Wx.planet: radius and gravity must be positive and omega not negative");
    auto o = std::make_unique<SphereObj>();
    o->model = std::make_unique<Sphere>(static_cast<int>(n), sp);
    return keep(spheres_, "Sphere", std::move(o));
  }
  // a Plane parameter operator: the new parameters reach the device model
  static Value set_plane(void *p, const std::function<void(PlaneParams &)> &edit) {
    PlaneObj &o = P(p);
    edit(o.params);
    o.model->set_params(o.params);
    return self_plane(p);
  }

  void register_all() {
    register_wx();
    register_sphere();
    register_plane();
  }

  // ---------------------------------------------------------------- Wx (cpp/main.cpp:108-142)
  void register_wx() {
    def("Wx", "sphere", {"int"}, "Sphere", [this](void *, Args &a) { return new_sphere(a[0].i, SphereParams{}); });
    def("Wx", "planet", {"int", "float", "float", "float"}, "Sphere", [this](void *, Args &a) {
      SphereParams sp;
      sp.radius = a[1].f;
      sp.omega = a[2].f;
      sp.gravity = a[3].f;
      return new_sphere(a[0].i, sp);
    });
    def("Wx", "plane", {"int", "int", "float", "float"}, "Plane", [this](void *, Args &a) {
      const int64_t nx = a[0].i, ny = a[1].i;
      need(nx >= 4 && ny >= 4 && nx <= 4000 && ny <= 4000 && nx * ny <= 8000000, "Wx.plane: need 4..4000 cells each way and at most 8 million cells");
      need(std::isfinite(a[2].f) && std::isfinite(a[3].f) && a[2].f > 0 && a[3].f > 0, "Wx.plane: domain sizes must be positive (m)");
      auto o = std::make_unique<PlaneObj>();
      o->params.nx = static_cast<int>(nx);
      o->params.ny = static_cast<int>(ny);
      o->params.lx = a[2].f;
      o->params.ly = a[3].f;
      o->model = std::make_unique<Plane>(o->params);
      return keep(planes_, "Plane", std::move(o));
    });
    def("Wx", "assert_close", {"float", "float", "float"}, "void", [](void *, Args &a) {
      const double x = a[0].f, y = a[1].f, tol = a[2].f;
      if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(tol) || tol < 0 || std::fabs(x - y) > tol)
        throw Error("assert_close failed: " + full_digits(x) + " vs " + full_digits(y) + " (tolerance " + full_digits(tol) + ")");
      return Value::none();
    });
    def("Wx", "assert_true", {"bool"}, "void", [](void *, Args &a) {
      need(a[0].b, "assert_true failed");
      return Value::none();
    });
    def("Wx", "print", {"string"}, "void", [](void *, Args &a) {
      std::cout << a[0].s << '\n';
      return Value::none();
    });
    def("Wx", "save_text", {"string", "string"}, "void", [](void *, Args &a) {
      write_text(a[0].s, a[1].s, "Wx.save_text");
      return Value::none();
    });
    def("Wx", "version", {}, "string", [](void *, Args &) { return Value::string("lmweather 0.1.4"); });
  }

  // ---------------------------------------------------------------- Sphere (cpp/main.cpp:144-384)
  void register_sphere() {
    auto real = [](double x) { return Value::real(x); };
    def("Sphere", "tilt", {"float"}, "Sphere", [](void *p, Args &a) {
      S(p).set_axis_tilt(finite(a[0].f, "tilt angle"));
      return self_sphere(p);
    });
    def("Sphere", "diffusion", {"float"}, "Sphere", [](void *p, Args &a) {
      need(std::isfinite(a[0].f) && a[0].f >= 0, "Sphere.diffusion: coefficient must be finite and not negative");
      S(p).set_diffusion(a[0].f);
      return self_sphere(p);
    });
    def("Sphere", "fields", {"string", "string", "string"}, "Sphere", [](void *p, Args &a) {
      S(p).fields(a[0].s, a[1].s, a[2].s);
      return self_sphere(p);
    });
    def("Sphere", "williamson2", {"float"}, "Sphere", [](void *p, Args &a) {
      S(p).setup(LMW_SETUP_WILLIAMSON2, a[0].f);
      return self_sphere(p);
    });
    def("Sphere", "williamson6", {}, "Sphere", [](void *p, Args &) {
      S(p).setup(LMW_SETUP_WILLIAMSON6);
      return self_sphere(p);
    });
    def("Sphere", "orography", {"string"}, "Sphere", [](void *p, Args &a) {
      S(p).orography(a[0].s);
      return self_sphere(p);
    });
    def("Sphere", "williamson5", {}, "Sphere", [](void *p, Args &) {
      S(p).setup(LMW_SETUP_WILLIAMSON5);
      return self_sphere(p);
    });
    def("Sphere", "tracer", {"string"}, "Sphere", [](void *p, Args &a) {
      S(p).tracer(a[0].s);
      return self_sphere(p);
    });
    def("Sphere", "advect_only", {}, "Sphere", [](void *p, Args &) {
      S(p).set_advect_only(true);
      return self_sphere(p);
    });
    def("Sphere", "williamson1", {"float"}, "Sphere", [](void *p, Args &a) {
      S(p).setup(LMW_SETUP_WILLIAMSON1, a[0].f);
      return self_sphere(p);
    });
    def("Sphere", "tracer_at", {"float", "float"}, "float", [real](void *p, Args &a) {
      return real(S(p).tracer_at(latitude(a[0].f, "Sphere.tracer_at"), finite(a[1].f, "Sphere.tracer_at: longitude") * kDeg));
    });
    def("Sphere", "tracer_mean", {}, "float", [real](void *p, Args &) { return real(S(p).tracer_mean()); });
    def("Sphere", "tracer_mass", {}, "float", [real](void *p, Args &) { return real(S(p).tracer_mass()); });
    def("Sphere", "tracer_min", {}, "float", [real](void *p, Args &) { return real(S(p).tracer_min()); });
    def("Sphere", "tracer_max", {}, "float", [real](void *p, Args &) { return real(S(p).tracer_max()); });
    def("Sphere", "tracer_error_l2", {}, "float", [real](void *p, Args &) { return real(S(p).tracer_error_l2()); });
    def("Sphere", "tracer_error_linf", {}, "float", [real](void *p, Args &) { return real(S(p).tracer_error_linf()); });
    def("Sphere", "mode", {"int", "int", "float", "float"}, "Sphere", [](void *p, Args &a) {
      positive(a[3].f, "Sphere.mode: mean depth must be positive");
      S(p).add_mode(narrow(a[0].i, "Sphere.mode"), narrow(a[1].i, "Sphere.mode"), finite(a[2].f, "amplitude"), a[3].f);
      return self_sphere(p);
    });
    def("Sphere", "perturb", {"float", "int", "int"}, "Sphere", [](void *p, Args &a) {
      need(a[1].i >= 0, "Sphere.perturb: seed must not be negative");
      S(p).perturb(finite(a[0].f, "amplitude"), static_cast<uint64_t>(a[1].i), narrow(a[2].i, "Sphere.perturb"));
      return self_sphere(p);
    });
    def("Sphere", "mark_reference", {}, "Sphere", [](void *p, Args &) {
      S(p).mark_reference();
      return self_sphere(p);
    });
    def("Sphere", "run", {"float"}, "Sphere", [](void *p, Args &a) {
      S(p).run(finite(a[0].f, "seconds"), 0);
      return self_sphere(p);
    });
    def("Sphere", "run_dt", {"float", "float"}, "Sphere", [](void *p, Args &a) {
      positive(a[1].f, "Sphere.run_dt: dt must be positive");
      S(p).run(finite(a[0].f, "seconds"), a[1].f);
      return self_sphere(p);
    });
    def("Sphere", "stable_dt", {}, "float", [real](void *p, Args &) { return real(S(p).stable_dt()); });
    def("Sphere", "time", {}, "float", [real](void *p, Args &) { return real(S(p).time()); });
    def("Sphere", "truncation", {}, "int", [](void *p, Args &) { return Value::integer(S(p).truncation()); });
    def("Sphere", "nlat", {}, "int", [](void *p, Args &) { return Value::integer(S(p).nlat()); });
    def("Sphere", "nlon", {}, "int", [](void *p, Args &) { return Value::integer(S(p).nlon()); });
    def("Sphere", "mean_height", {}, "float", [real](void *p, Args &) { return real(S(p).mean_height()); });
    def("Sphere", "energy", {}, "float", [real](void *p, Args &) { return real(S(p).energy()); });
    def("Sphere", "kinetic_energy", {}, "float", [real](void *p, Args &) { return real(S(p).kinetic_energy()); });
    def("Sphere", "potential_energy", {}, "float", [real](void *p, Args &) { return real(S(p).potential_energy()); });
    def("Sphere", "save_spectrum", {"string"}, "void", [](void *p, Args &a) {
      S(p).save_spectrum(a[0].s);
      return Value::none();
    });
    def("Sphere", "enstrophy", {}, "float", [real](void *p, Args &) { return real(S(p).potential_enstrophy()); });
    def("Sphere", "max_wind", {}, "float", [real](void *p, Args &) { return real(S(p).max_wind()); });
    def("Sphere", "error_l2", {}, "float", [real](void *p, Args &) { return real(S(p).height_error_l2()); });
    def("Sphere", "error_linf", {}, "float", [real](void *p, Args &) { return real(S(p).height_error_linf()); });
    def("Sphere", "height_at", {"float", "float"}, "float", [real](void *p, Args &a) {
      return real(S(p).height_at(latitude(a[0].f, "Sphere.height_at"), finite(a[1].f, "Sphere.height_at: longitude") * kDeg));
    });
    // the winds are singular at the poles (u = U / cos(lat)): open interval
    def("Sphere", "u_at", {"float", "float"}, "float", [real](void *p, Args &a) {
      need(std::fabs(a[0].f) < 90.0, "Sphere.u_at: latitude must be inside (-90, 90) degrees");
      return real(S(p).wind_at(a[0].f * kDeg, finite(a[1].f, "Sphere.u_at: longitude") * kDeg, true));
    });
    def("Sphere", "v_at", {"float", "float"}, "float", [real](void *p, Args &a) {
      need(std::fabs(a[0].f) < 90.0, "Sphere.v_at: latitude must be inside (-90, 90) degrees");
      return real(S(p).wind_at(a[0].f * kDeg, finite(a[1].f, "Sphere.v_at: longitude") * kDeg, false));
    });
    def("Sphere", "vorticity_at", {"float", "float"}, "float", [real](void *p, Args &a) {
      return real(S(p).vorticity_at(latitude(a[0].f, "Sphere.vorticity_at"), finite(a[1].f, "Sphere.vorticity_at: longitude") * kDeg));
    });
    def("Sphere", "mode_coefficient", {"int", "int", "string"}, "float", [real](void *p, Args &a) {
      need(a[2].s == "re" || a[2].s == "im", "Sphere.mode_coefficient: part must be re or im");
      return real(S(p).mode_coefficient(narrow(a[0].i, "Sphere.mode_coefficient"), narrow(a[1].i, "Sphere.mode_coefficient"), a[2].s == "im"));
    });
    def("Sphere", "wave_crest", {"int", "float"}, "float", [real](void *p, Args &a) {
      return real(S(p).wave_crest(narrow(a[0].i, "Sphere.wave_crest"), latitude(a[1].f, "Sphere.wave_crest")) / kDeg);
    });
    def("Sphere", "save_csv", {"string"}, "void", [](void *p, Args &a) {
      S(p).save_csv(a[0].s);
      return Value::none();
    });
    def("Sphere", "save_zonal_mean", {"string"}, "void", [](void *p, Args &a) {
      S(p).save_zonal_mean(a[0].s);
      return Value::none();
    });
  }

  // ---------------------------------------------------------------- Plane (cpp/main.cpp:386-494)
  void register_plane() {
    auto real = [](double x) { return Value::real(x); };
    def("Plane", "gravity", {"float"}, "Plane", [](void *p, Args &a) {
      positive(a[0].f, "Plane.gravity: must be positive");
      return set_plane(p, [&](PlaneParams &q) { q.gravity = a[0].f; });
    });
    def("Plane", "depth", {"float"}, "Plane", [](void *p, Args &a) {
      positive(a[0].f, "Plane.depth: must be positive (m)");
      need(!P(p).started, "Plane.depth: set the depth before fields() or perturb()");
      set_plane(p, [&](PlaneParams &q) { q.depth = a[0].f; });
      P(p).model->reset();  // the resting layer of the new depth
      return self_plane(p);
    });
    def("Plane", "coriolis", {"float", "float"}, "Plane", [](void *p, Args &a) {
      const double f0 = finite(a[0].f, "f0"), beta = finite(a[1].f, "beta");
      return set_plane(p, [&](PlaneParams &q) {
        q.f0 = f0;
        q.beta = beta;
      });
    });
    def("Plane", "linear", {}, "Plane", [](void *p, Args &) { return set_plane(p, [](PlaneParams &q) { q.linear = true; }); });
    def("Plane", "nonlinear", {}, "Plane", [](void *p, Args &) { return set_plane(p, [](PlaneParams &q) { q.linear = false; }); });
    def("Plane", "viscosity", {"float"}, "Plane", [](void *p, Args &a) {
      need(std::isfinite(a[0].f) && a[0].f >= 0, "Plane.viscosity: must be finite and not negative");
      return set_plane(p, [&](PlaneParams &q) { q.viscosity = a[0].f; });
    });
    def("Plane", "drag", {"float"}, "Plane", [](void *p, Args &a) {
      need(std::isfinite(a[0].f) && a[0].f >= 0, "Plane.drag: must be finite and not negative");
      return set_plane(p, [&](PlaneParams &q) { q.drag = a[0].f; });
    });
    def("Plane", "fields", {"string", "string", "string"}, "Plane", [](void *p, Args &a) {
      P(p).model->fields(a[0].s, a[1].s, a[2].s);
      P(p).started = true;
      return self_plane(p);
    });
    def("Plane", "perturb", {"float", "int"}, "Plane", [](void *p, Args &a) {
      need(a[1].i >= 0, "Plane.perturb: seed must not be negative");
      P(p).model->perturb(finite(a[0].f, "amplitude"), static_cast<uint64_t>(a[1].i));
      P(p).started = true;
      return self_plane(p);
    });
    def("Plane", "run", {"float"}, "Plane", [](void *p, Args &a) {
      P(p).model->run(finite(a[0].f, "seconds"), 0);
      P(p).started = true;
      return self_plane(p);
    });
    def("Plane", "run_dt", {"float", "float"}, "Plane", [](void *p, Args &a) {
      positive(a[1].f, "Plane.run_dt: dt must be positive");
      P(p).model->run(finite(a[0].f, "seconds"), a[1].f);
      P(p).started = true;
      return self_plane(p);
    });
    def("Plane", "stable_dt", {}, "float", [real](void *p, Args &) { return real(P(p).model->stable_dt()); });
    def("Plane", "time", {}, "float", [real](void *p, Args &) { return real(P(p).model->time()); });
    def("Plane", "wave_speed", {}, "float", [real](void *p, Args &) { return real(std::sqrt(P(p).params.gravity * P(p).params.depth)); });
    def("Plane", "rossby_radius", {}, "float", [real](void *p, Args &) {
      need(P(p).params.f0 != 0, "Plane.rossby_radius: f0 is zero");
      return real(std::sqrt(P(p).params.gravity * P(p).params.depth) / std::fabs(P(p).params.f0));
    });
    def("Plane", "mean_depth", {}, "float", [real](void *p, Args &) { return real(P(p).model->mean_depth()); });
    def("Plane", "energy", {}, "float", [real](void *p, Args &) { return real(P(p).model->energy()); });
    def("Plane", "max_speed", {}, "float", [real](void *p, Args &) { return real(P(p).model->max_speed()); });
    def("Plane", "pv_anomaly_rms", {}, "float", [real](void *p, Args &) { return real(P(p).model->potential_vorticity_anomaly_rms()); });
    auto inside = [](void *p, double x, double y, const std::string &who) {
      need(x >= 0 && x <= P(p).params.lx && y >= 0 && y <= P(p).params.ly, who + ": point outside the domain");
    };
    def("Plane", "h_at", {"float", "float"}, "float", [real, inside](void *p, Args &a) {
      inside(p, a[0].f, a[1].f, "Plane.h_at");
      return real(P(p).model->h_at(a[0].f, a[1].f));
    });
    def("Plane", "u_at", {"float", "float"}, "float", [real, inside](void *p, Args &a) {
      inside(p, a[0].f, a[1].f, "Plane.u_at");
      return real(P(p).model->u_at(a[0].f, a[1].f));
    });
    def("Plane", "v_at", {"float", "float"}, "float", [real, inside](void *p, Args &a) {
      inside(p, a[0].f, a[1].f, "Plane.v_at");
      return real(P(p).model->v_at(a[0].f, a[1].f));
    });
    def("Plane", "zonal_amplitude", {"int", "int"}, "float", [real](void *p, Args &a) {
      return real(P(p).model->zonal_amplitude(narrow(a[0].i, "Plane.zonal_amplitude"), narrow(a[1].i, "Plane.zonal_amplitude")));
    });
    def("Plane", "zonal_crest", {"int", "int"}, "float", [real](void *p, Args &a) {
      return real(P(p).model->zonal_crest(narrow(a[0].i, "Plane.zonal_crest"), narrow(a[1].i, "Plane.zonal_crest")));
    });
    def("Plane", "harmonic_energy", {"int"}, "float", [real](void *p, Args &a) {
      return real(P(p).model->harmonic_energy(narrow(a[0].i, "Plane.harmonic_energy")));
    });
    def("Plane", "v_zonal_crest", {"int", "int"}, "float", [real](void *p, Args &a) {
      return real(P(p).model->v_zonal_crest(narrow(a[0].i, "Plane.v_zonal_crest"), narrow(a[1].i, "Plane.v_zonal_crest")));
    });
    def("Plane", "save_csv", {"string"}, "void", [](void *p, Args &a) {
      P(p).model->save_csv(a[0].s);
      return Value::none();
    });
    def("Plane", "save_row", {"int", "string"}, "void", [](void *p, Args &a) {
      need(a[0].i >= 0 && a[0].i < P(p).params.ny, "Plane.save_row: row outside the grid");
      P(p).model->save_row(static_cast<int>(a[0].i), a[1].s);
      return Value::none();
    });
    def("Plane", "save_column", {"int", "string"}, "void", [](void *p, Args &a) {
      need(a[0].i >= 0 && a[0].i < P(p).params.nx, "Plane.save_column: column outside the grid");
      P(p).model->save_column(static_cast<int>(a[0].i), a[1].s);
      return Value::none();
    });
  }
};

}  // namespace ops
}  // namespace lmw
