// Expression grids: the initial-condition operators that take formula strings (Sphere.fields / orography / tracer
// over lon, lat and Plane.fields over x, y; cpp/main.cpp:158-172,212-228,252-268,410-431) evaluated on the GPU.
//
// The host compiles a string once into postfix code for a small stack machine (compile_expr); every grid point is
// one GPU thread that runs the code with its own coordinates.  All threads run the same instructions, so the
// instruction stream is uniform (scalar loads) and there is no divergence.  The grammar, the names, the functions
// and the error texts are those of the reference's expression compiler (cpp/expr.hpp:16-166): a caller's strings
// parse, fail and evaluate the same way.
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "common.h"

namespace lmw {

enum ExprOp : int {
  XO_CONST, XO_VAR0, XO_VAR1,
  XO_NEG, XO_SIN, XO_COS, XO_TAN, XO_ASIN, XO_ACOS, XO_ATAN, XO_SINH, XO_COSH, XO_TANH, XO_EXP, XO_LOG, XO_LOG10,
  XO_SQRT, XO_ABS, XO_FLOOR, XO_CEIL, XO_SIGN,
  XO_ADD, XO_SUB, XO_MUL, XO_DIV, XO_POW, XO_ATAN2, XO_MIN, XO_MAX, XO_MOD
};

struct ExprIns {
  int op;
  int pad;
  double c;  // the value of XO_CONST
};

// Postfix code of one expression and the evaluation stack it needs (values held at once).
struct ExprProgram {
  std::vector<ExprIns> code;
  int depth = 0;
};

// Deepest evaluation stack of the device interpreter.  The reference's nesting limit (200 levels of sums and signs,
// cpp/expr.hpp:51-58) keeps every accepted expression well below it; compile_expr checks.
constexpr int kExprMaxDepth = 256;
constexpr int kExprSmallDepth = 16;  // the common case: a stack that stays in registers

// Compiles text over the variables vars[0], vars[1] and the named constants params (looked up after the variables,
// before pi and e).  Throws the reference's std::runtime_error texts: 'expression "<text>": <what> at column <c>'.
ExprProgram compile_expr(const std::string &text, const char *var0, const char *var1,
                         const std::vector<std::pair<std::string, double>> &params);

// One program's code on the device: the instructions of up to three expressions back to back.
struct ExprSet {
  const ExprIns *code = nullptr;
  int start[3] = {0, 0, 0}, len[3] = {0, 0, 0};
};

// Runs one program at (v0, v1).  D bounds the stack: the top of the stack is a register, the rest st[0 .. D-1].  The
// arithmetic is the reference's operation for operation (one IEEE operation per node, no contraction across
// nodes); the transcendental functions are the device library's.
template <int D> __device__ inline double expr_eval(const ExprIns *__restrict__ p, int n, double v0, double v1) {
#pragma clang fp contract(off)
  double st[D];
  double top = 0;
  int sp = 0;
  for (int i = 0; i < n; i++) {
    const int op = p[i].op;
    if (op <= XO_VAR1) {  // push
      st[sp] = top;
      sp++;
      top = op == XO_CONST ? p[i].c : op == XO_VAR0 ? v0 : v1;
      continue;
    }
    if (op < XO_ADD) {  // one argument
      switch (op) {
        case XO_NEG: top = -top; break;
        case XO_SIN: top = sin(top); break;
        case XO_COS: top = cos(top); break;
        case XO_TAN: top = tan(top); break;
        case XO_ASIN: top = asin(top); break;
        case XO_ACOS: top = acos(top); break;
        case XO_ATAN: top = atan(top); break;
        case XO_SINH: top = sinh(top); break;
        case XO_COSH: top = cosh(top); break;
        case XO_TANH: top = tanh(top); break;
        case XO_EXP: top = exp(top); break;
        case XO_LOG: top = log(top); break;
        case XO_LOG10: top = log10(top); break;
        case XO_SQRT: top = sqrt(top); break;
        case XO_ABS: top = fabs(top); break;
        case XO_FLOOR: top = floor(top); break;
        case XO_CEIL: top = ceil(top); break;
        default: top = static_cast<double>((top > 0) - (top < 0)); break;  // XO_SIGN
      }
      continue;
    }
    sp--;
    const double a = st[sp], b = top;  // a (op) b
    switch (op) {
      case XO_ADD: top = a + b; break;
      case XO_SUB: top = a - b; break;
      case XO_MUL: top = a * b; break;
      case XO_DIV: top = a / b; break;
      case XO_POW: top = pow(a, b); break;
      case XO_ATAN2: top = atan2(a, b); break;
      case XO_MIN: top = a < b ? a : b; break;
      case XO_MAX: top = a > b ? a : b; break;
      default: top = fmod(a, b); break;  // XO_MOD
    }
  }
  return top;
}

// The device copy of up to three compiled programs and a flag word, owned by a model.
struct ExprDevice {
  DevArray<ExprIns> code;
  DevArray<int> bad;  // [3]: a non-finite value of expression (or sum) e was seen
  int depth = 0;
  ExprSet load(const ExprProgram *const progs[], int count, hipStream_t stream);
};

}  // namespace lmw
