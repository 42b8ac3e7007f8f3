// Expression grids on the GPU: the compiler of formula strings into postfix code (the reference's grammar and error
// texts, cpp/expr.hpp:16-166), the evaluation kernel of the sphere's fields / orography / tracer operators
// (cpp/main.cpp:158-172,212-228,252-268) and the SphereCore methods behind them.  The plane's evaluation kernel is
// in plane_core.h.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>

#include "expr.h"
#include "sphere_core.h"

namespace lmw {

namespace {

constexpr int kMaxNest = 200;    // cpp/expr.hpp:51: nesting of sums and signs
constexpr int kMaxLeaves = 2000;  // cpp/expr.hpp:52: atoms (numbers, names, calls, parentheses)

struct FnEntry {
  const char *name;
  int op, args;
};
// cpp/expr.hpp:131-164
const FnEntry kFunctions[] = {
    {"sin", XO_SIN, 1},   {"cos", XO_COS, 1},     {"tan", XO_TAN, 1},     {"asin", XO_ASIN, 1},   {"acos", XO_ACOS, 1},
    {"atan", XO_ATAN, 1}, {"sinh", XO_SINH, 1},   {"cosh", XO_COSH, 1},   {"tanh", XO_TANH, 1},   {"exp", XO_EXP, 1},
    {"log", XO_LOG, 1},   {"ln", XO_LOG, 1},      {"log10", XO_LOG10, 1}, {"sqrt", XO_SQRT, 1},   {"abs", XO_ABS, 1},
    {"floor", XO_FLOOR, 1}, {"ceil", XO_CEIL, 1}, {"sign", XO_SIGN, 1},   {"atan2", XO_ATAN2, 2}, {"pow", XO_POW, 2},
    {"min", XO_MIN, 2},   {"max", XO_MAX, 2},     {"mod", XO_MOD, 2},
};

// Recursive descent over sums, products, signs, powers and atoms, emitting postfix code as it goes.  The nesting
// and leaf counters sit where the reference's do, so an over-deep or over-long string fails at the same column.
class Compiler {
 public:
  Compiler(const std::string &text, const char *v0, const char *v1, const std::vector<std::pair<std::string, double>> &params)
      : src_(text), v0_(v0), v1_(v1), params_(params) {}

  ExprProgram run() {
    sum();
    skip();
    if (pos_ != src_.size()) fail(std::string("unexpected '") + src_[pos_] + "'");
    if (prog_.depth > kExprMaxDepth) fail("expression needs a deeper evaluation stack than the device provides");
    return std::move(prog_);
  }

 private:
  const std::string &src_;
  const char *v0_, *v1_;
  const std::vector<std::pair<std::string, double>> &params_;
  size_t pos_ = 0;
  int nest_ = 0, leaves_ = 0, held_ = 0;
  ExprProgram prog_;

  struct Nest {
    Compiler &c;
    explicit Nest(Compiler &x) : c(x) {
      if (++c.nest_ > kMaxNest) c.fail("expression is nested too deeply");
    }
    ~Nest() { --c.nest_; }
  };

  [[noreturn]] void fail(const std::string &what) const {
    const std::string shown = src_.size() <= 120 ? src_ : src_.substr(0, 117) + "...";
    throw Error(LMW_ERUNTIME, "expression \"" + shown + "\": " + what + " at column " + std::to_string(pos_ + 1));
  }
  void emit(int op, double c = 0) {
    prog_.code.push_back(ExprIns{op, 0, c});
    if (op <= XO_VAR1) prog_.depth = std::max(prog_.depth, ++held_);
    else if (op >= XO_ADD) held_--;
  }
  void skip() {
    while (pos_ < src_.size() && std::isspace(static_cast<unsigned char>(src_[pos_]))) pos_++;
  }
  bool eat(char c) {
    skip();
    if (pos_ < src_.size() && src_[pos_] == c) {
      pos_++;
      return true;
    }
    return false;
  }
  bool eat_starstar() {
    skip();
    if (pos_ + 1 < src_.size() && src_[pos_] == '*' && src_[pos_ + 1] == '*') {
      pos_ += 2;
      return true;
    }
    return false;
  }
  void sum() {  // terms joined by + and -, left to right
    Nest guard(*this);
    product();
    for (;;) {
      if (eat('+')) {
        product();
        emit(XO_ADD);
      } else if (eat('-')) {
        product();
        emit(XO_SUB);
      } else {
        return;
      }
    }
  }
  void product() {  // signed factors joined by * and /, left to right
    signed_factor();
    for (;;) {
      if (eat('*')) {
        signed_factor();
        emit(XO_MUL);
      } else if (eat('/')) {
        signed_factor();
        emit(XO_DIV);
      } else {
        return;
      }
    }
  }
  void signed_factor() {  // any number of leading signs, then a power
    Nest guard(*this);
    if (eat('-')) {
      signed_factor();
      emit(XO_NEG);
      return;
    }
    if (eat('+')) {
      signed_factor();
      return;
    }
    power();
  }
  void power() {  // atom ^ signed factor, right associative (the exponent may carry a sign); ** is ^
    atom();
    if (eat('^') || eat_starstar()) {
      signed_factor();
      emit(XO_POW);
    }
  }
  void atom() {
    skip();
    if (pos_ >= src_.size()) fail("unexpected end");
    const char c = src_[pos_];
    if (++leaves_ > kMaxLeaves) fail("expression has too many terms");
    if (c == '(') {
      pos_++;
      sum();
      if (!eat(')')) fail("expected ')'");
      return;
    }
    if (std::isdigit(static_cast<unsigned char>(c)) || c == '.') {
      size_t used = 0;
      double x = 0;
      try {
        x = std::stod(src_.substr(pos_), &used);  // the C library's strtod, as the reference reads numbers
      } catch (const std::exception &) {
        fail("number is malformed or out of range");
      }
      pos_ += used;
      emit(XO_CONST, x);
      return;
    }
    if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
      const size_t s = pos_;
      while (pos_ < src_.size() && (std::isalnum(static_cast<unsigned char>(src_[pos_])) || src_[pos_] == '_')) pos_++;
      const std::string id = src_.substr(s, pos_ - s);
      skip();
      if (pos_ < src_.size() && src_[pos_] == '(') {
        pos_++;
        int args = 0;
        if (!eat(')')) {
          do {
            sum();
            args++;
          } while (eat(','));
          if (!eat(')')) fail("expected ')'");
        }
        call(id, args);
        return;
      }
      if (id == v0_) return emit(XO_VAR0);
      if (id == v1_) return emit(XO_VAR1);
      for (const auto &p : params_)
        if (p.first == id) return emit(XO_CONST, p.second);
      if (id == "pi") return emit(XO_CONST, 3.14159265358979323846);
      if (id == "e") return emit(XO_CONST, 2.71828182845904523536);
      fail("unknown name '" + id + "'");
    }
    fail(std::string("unexpected '") + c + "'");
  }
  void call(const std::string &f, int args) {
    for (const FnEntry &e : kFunctions) {
      if (f != e.name) continue;
      if (args != e.args) fail(f + (e.args == 1 ? " takes 1 argument" : " takes 2 arguments"));
      emit(e.op);
      return;
    }
    fail("unknown function '" + f + "'");
  }
};

// The sphere's expressions at every Gauss grid point: lonlat = lon[K] then lat[J].  nexpr 3 (h, u, v: the sum is
// checked, as cpp/main.cpp:170 does) or 1 (orography or tracer).  Writes out[bl][e][j][k] for bl < nb, the layout
// of the analysis staging buffer.
template <int D>
__global__ __launch_bounds__(256) void k_expr_sphere(ExprSet X, int nexpr, const double *__restrict__ lonlat, int J, int K,
                                                     double *__restrict__ out, int nb, int *__restrict__ bad) {
  const int k = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (k >= K) return;
  const double lon = lonlat[k], lat = lonlat[K + j];
  double v[3] = {0, 0, 0};
#pragma unroll
  for (int e = 0; e < 3; e++)
    if (e < nexpr) v[e] = expr_eval<D>(X.code + X.start[e], X.len[e], lon, lat);
  const double check = nexpr == 3 ? (v[0] + v[1]) + v[2] : v[0];
  if (!isfinite(check)) bad[0] = 1;
  const size_t plane = static_cast<size_t>(J) * K, at = static_cast<size_t>(j) * K + k;
  for (int bl = 0; bl < nb; bl++)
#pragma unroll
    for (int e = 0; e < 3; e++)
      if (e < nexpr) out[(static_cast<size_t>(bl) * 3 + e) * plane + at] = v[e];
}

std::vector<std::pair<std::string, double>> sphere_params(const lmw_sphere_params &p) {
  return {{"a", p.radius}, {"omega", p.omega}, {"g", p.gravity}};  // cpp/main.cpp:162
}

}  // namespace

ExprProgram compile_expr(const std::string &text, const char *var0, const char *var1,
                         const std::vector<std::pair<std::string, double>> &params) {
  return Compiler(text, var0, var1, params).run();
}

ExprSet ExprDevice::load(const ExprProgram *const progs[], int count, hipStream_t stream) {
  ExprSet X;
  size_t total = 0;
  depth = 0;
  for (int e = 0; e < count; e++) {
    X.start[e] = static_cast<int>(total);
    X.len[e] = static_cast<int>(progs[e]->code.size());
    total += progs[e]->code.size();
    depth = std::max(depth, progs[e]->depth);
  }
  if (code.n < total) code.alloc(total);
  for (int e = 0; e < count; e++) code.upload(progs[e]->code.data(), progs[e]->code.size(), X.start[e], stream);
  if (bad.n < 3) bad.alloc(3);
  LMW_HIP(hipMemsetAsync(bad.p, 0, 3 * sizeof(int), stream));
  X.code = code.p;
  return X;
}

// lon (sphere.cpp:121) and lat (sphere.cpp:120) of the grid on the host, with the host's libm as the reference
// computes them, then on the device: [K] longitudes, [J] latitudes
void SphereCore::upload_lonlat() {
  std::vector<double> mh(J), ll(static_cast<size_t>(K) + J);
  mu.download(mh.data(), J, 0, stream);
  sync();
  constexpr double kPi = 3.14159265358979323846;
  for (int k = 0; k < K; k++) ll[k] = 2 * kPi * k / K;
  for (int j = 0; j < J; j++) ll[K + j] = std::asin(mh[j]);
  if (lonlat.n < ll.size()) lonlat.alloc(ll.size());
  lonlat.upload(ll.data(), ll.size(), 0, stream);
}

// Evaluates count (1 or 3) compiled expressions at the grid into in3 for nb members and returns whether every
// value (or the sum of the three) is finite.
bool SphereCore::eval_exprs(const ExprProgram *const progs[], int count, int nb) {
  const ExprSet X = expr.load(progs, count, stream);
  upload_lonlat();
  ensure_in3(nb);
  const dim3 grd((K + 255) / 256, J);
  if (expr.depth <= kExprSmallDepth)
    hipLaunchKernelGGL(k_expr_sphere<kExprSmallDepth>, grd, dim3(256), 0, stream, X, count, lonlat.p, J, K, in3.p, nb, expr.bad.p);
  else
    hipLaunchKernelGGL(k_expr_sphere<kExprMaxDepth>, grd, dim3(256), 0, stream, X, count, lonlat.p, J, K, in3.p, nb, expr.bad.p);
  LMW_LAUNCH_CHECK();
  int bad = 0;
  expr.bad.download(&bad, 1, 0, stream);
  sync();
  return bad == 0;
}

// Sphere.fields (cpp/main.cpp:158-172) for members b0 .. b0 + nb - 1: h, u, v from expressions in lon, lat and the
// planet's a, omega, g, then set_state_grid.  The grid is evaluated once; every chunk of members is analysed from
// it.  A non-finite value leaves the state unchanged.
void SphereCore::fields_expr(int b0, int nb, const std::string &h, const std::string &u, const std::string &v) {
  const auto par = sphere_params(prm);
  const ExprProgram ph = compile_expr(h, "lon", "lat", par), pu = compile_expr(u, "lon", "lat", par), pv = compile_expr(v, "lon", "lat", par);
  const ExprProgram *progs[3] = {&ph, &pu, &pv};
  if (!eval_exprs(progs, 3, std::min(nb, kSetupChunk))) runtime("Sphere.fields: an expression is not finite on the grid");
  for (int c0 = 0; c0 < nb; c0 += kSetupChunk) {
    state_from_in3(b0 + c0, std::min(kSetupChunk, nb - c0));
    sync();
  }
}

// Sphere.orography (cpp/main.cpp:212-228) for members b0 .. b0 + nb - 1: one grid, one shared plane
void SphereCore::orography_expr(int b0, int nb, const std::string &text) {
  const ExprProgram p = compile_expr(text, "lon", "lat", sphere_params(prm));
  const ExprProgram *progs[1] = {&p};
  if (!eval_exprs(progs, 1, 1)) runtime("Sphere.orography: the expression is not finite on the grid");
  const size_t plane = static_cast<size_t>(J) * K;
  std::vector<double> host(plane);
  in3.download(host.data(), plane, 0, stream);
  sync();
  set_orography_run(b0, nb, host.data(), in3.p);
}

// Sphere.tracer (cpp/main.cpp:252-268)
void SphereCore::tracer_expr(const std::string &text) {
  if (B != 1) invalid("a tracer needs a single-member sphere");
  const ExprProgram p = compile_expr(text, "lon", "lat", sphere_params(prm));
  const ExprProgram *progs[1] = {&p};
  if (!eval_exprs(progs, 1, 1)) runtime("Sphere.tracer: the expression is not finite on the grid");
  tracer_from_in3();
}

}  // namespace lmw
