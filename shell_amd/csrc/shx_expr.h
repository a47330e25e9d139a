/*
 * shx_expr.h -- the time functions of the reference's config.cfg as small
 * compiled programs: an expression of the muparser dialect the reference
 * reads through Function (Function.h:7-51, MAKE_FUNCTION macros.h:56-68) is
 * parsed on the host into a postfix program over the variable t, which the
 * host and the GPU kernels evaluate with a short value stack.
 *
 * Grammar (muparser's operator set and precedence, lowest first):
 *   ?:  ||  &&  < > <= >= == !=  + -  * /  unary - +  ^ (right associative)
 * with numbers, the variable t, the constants pi, _pi, _e, and the
 * functions sin cos tan asin acos atan sinh cosh tanh asinh acosh atanh
 * exp ln log (base 10) log10 log2 sqrt abs sign rint, min(a,b), max(a,b).
 * The position variables x, y, z (MAKE_FUNCTION, macros.h:56-58) are
 * accepted where the caller allows them: the growth factors, which the
 * reference evaluates at each element's reference point (ShellSystem.C:
 * 2401-2404).  The density d is refused.
 */
#ifndef SHX_EXPR_H
#define SHX_EXPR_H

#include <math.h>

#if defined(__HIPCC__)
#define SHX_EXPR_HD __host__ __device__ inline
#else
#define SHX_EXPR_HD inline
#endif

namespace shx {

constexpr int kExprOps = 96;      /* postfix operations of one program */
constexpr int kExprConsts = 24;   /* numeric literals of one program */
constexpr int kExprStack = 16;    /* evaluation stack depth */

enum ExprOp : unsigned char {
  EX_END = 0, EX_CONST, EX_T, EX_X, EX_Y, EX_Z,
  EX_ADD, EX_SUB, EX_MUL, EX_DIV, EX_POW, EX_NEG,
  EX_LT, EX_GT, EX_LE, EX_GE, EX_EQ, EX_NE, EX_AND, EX_OR, EX_SEL,
  EX_SIN, EX_COS, EX_TAN, EX_ASIN, EX_ACOS, EX_ATAN, EX_SINH, EX_COSH, EX_TANH,
  EX_ASINH, EX_ACOSH, EX_ATANH, EX_EXP, EX_LN, EX_LOG10, EX_LOG2, EX_SQRT, EX_ABS,
  EX_SIGN, EX_RINT, EX_MIN, EX_MAX
};

/* a compiled expression; n == 0: empty (the caller's default applies) */
struct ExprProg {
  int n;                          /* operations */
  int uses_t;                     /* 0: constant in time */
  int uses_xyz;                   /* 1: depends on the position (x, y, z) */
  unsigned char op[kExprOps];
  unsigned char arg[kExprOps];    /* EX_CONST: index into k */
  double k[kExprConsts];
};

/* value at time t and position (x, y, z) */
SHX_EXPR_HD double expr_eval_at(const ExprProg& p, double t, double x, double y, double z)
{
  double s[kExprStack];
  int sp = 0;
  for (int i = 0; i < p.n; ++i) {
    const unsigned char o = p.op[i];
    if (o == EX_CONST) { s[sp++] = p.k[p.arg[i]]; continue; }
    if (o == EX_T) { s[sp++] = t; continue; }
    if (o == EX_X) { s[sp++] = x; continue; }
    if (o == EX_Y) { s[sp++] = y; continue; }
    if (o == EX_Z) { s[sp++] = z; continue; }
    if (o == EX_SEL) {            /* c ? a : b, both branches evaluated (pure functions) */
      const double b = s[--sp], a = s[--sp], c = s[--sp];
      s[sp++] = c != 0 ? a : b;
      continue;
    }
    if (o >= EX_ADD && o <= EX_OR && o != EX_NEG) {
      const double b = s[--sp], a = s[--sp];
      double r = 0;
      switch (o) {
        case EX_ADD: r = a + b; break;
        case EX_SUB: r = a - b; break;
        case EX_MUL: r = a * b; break;
        case EX_DIV: r = a / b; break;
        case EX_POW: r = pow(a, b); break;
        case EX_LT: r = a < b; break;
        case EX_GT: r = a > b; break;
        case EX_LE: r = a <= b; break;
        case EX_GE: r = a >= b; break;
        case EX_EQ: r = a == b; break;
        case EX_NE: r = a != b; break;
        case EX_AND: r = (a != 0) && (b != 0); break;
        case EX_OR: r = (a != 0) || (b != 0); break;
        default: break;
      }
      s[sp++] = r;
      continue;
    }
    if (o == EX_MIN || o == EX_MAX) {
      const double b = s[--sp], a = s[--sp];
      s[sp++] = o == EX_MIN ? fmin(a, b) : fmax(a, b);
      continue;
    }
    double& a = s[sp - 1];
    switch (o) {
      case EX_NEG: a = -a; break;
      case EX_SIN: a = sin(a); break;
      case EX_COS: a = cos(a); break;
      case EX_TAN: a = tan(a); break;
      case EX_ASIN: a = asin(a); break;
      case EX_ACOS: a = acos(a); break;
      case EX_ATAN: a = atan(a); break;
      case EX_SINH: a = sinh(a); break;
      case EX_COSH: a = cosh(a); break;
      case EX_TANH: a = tanh(a); break;
      case EX_ASINH: a = asinh(a); break;
      case EX_ACOSH: a = acosh(a); break;
      case EX_ATANH: a = atanh(a); break;
      case EX_EXP: a = exp(a); break;
      case EX_LN: a = log(a); break;
      case EX_LOG10: a = log10(a); break;
      case EX_LOG2: a = log2(a); break;
      case EX_SQRT: a = sqrt(a); break;
      case EX_ABS: a = fabs(a); break;
      case EX_SIGN: a = (a > 0) ? 1.0 : (a < 0 ? -1.0 : 0.0); break;
      case EX_RINT: a = rint(a); break;
      default: break;
    }
  }
  return sp > 0 ? s[sp - 1] : 0.0;
}

SHX_EXPR_HD double expr_eval(const ExprProg& p, double t) { return expr_eval_at(p, t, 0.0, 0.0, 0.0); }

/* host: compile `text`; false with a message in err (errlen bytes) on a
 * syntax error, an unknown name, a variable other than t (and x, y, z when
 * positions is true), or a program beyond the fixed capacity */
bool expr_compile(const char* text, ExprProg& out, char* err, int errlen, bool positions = false);

} // namespace shx
#endif
