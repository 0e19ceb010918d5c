// CPU ORACLE — test infrastructure only.
//
// A C++ restatement of the reference's CPU BytecodeTape sweeps, used by the
// tests, by the package's smoke check and by bench.py's cpu_baseline leg as
// the checker.  The product (echidna_amd/, libechidna_hip.so) never links,
// loads or calls anything in this directory.
//
// Followed (reference file:line):
//   opcode semantics        src/opcode.rs:228-318 (eval_forward),
//                           src/opcode.rs:331-511 (reverse_partials, powi)
//   derivative helpers      src/kernels/mod.rs:85-285
//   forward sweep           src/bytecode_tape/forward.rs:11-43, 194-229
//   reverse sweep           src/bytecode_tape/reverse.rs:32-105, 138-145
//   forward-over-reverse    src/bytecode_tape/tangent.rs:97-147, 191-262, 268-310
//   Dual arithmetic         src/dual.rs:44-460, src/traits/std_ops.rs:14-83
//   IsAllZero               src/float.rs:65-98
//   Taylor<F,3>             src/taylor.rs:49-370, src/taylor_ops.rs:18-916,
//                           src/traits/taylor_std_ops.rs:15-95
//   kinks                   src/bytecode_tape/forward.rs:87-187
//
// Every function here is templated the way the reference's generic code is
// (over f32 / f64 / Dual<F> / Taylor<F,3>), and evaluates in exactly the
// order the reference writes its expressions.  Compile with
// -ffp-contract=off -fno-fast-math: Rust never contracts to FMA.
//
// Bit-level parity is unpinned where the reference calls Rust-std code that
// is not libm (asinh/acosh/atanh identities, powi via repeated squaring) —
// those are restated from their published formulas and pinned by tolerance
// tests and closed forms only.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace orc {

enum Op : uint32_t {
  OP_INPUT = 0, OP_CONST = 1, OP_ADD = 2, OP_SUB = 3, OP_MUL = 4, OP_DIV = 5,
  OP_REM = 6, OP_POWF = 7, OP_ATAN2 = 8, OP_HYPOT = 9, OP_MAX = 10, OP_MIN = 11,
  OP_NEG = 12, OP_RECIP = 13, OP_SQRT = 14, OP_CBRT = 15, OP_POWI = 16,
  OP_EXP = 17, OP_EXP2 = 18, OP_EXPM1 = 19, OP_LN = 20, OP_LOG2 = 21,
  OP_LOG10 = 22, OP_LN1P = 23, OP_SIN = 24, OP_COS = 25, OP_TAN = 26,
  OP_ASIN = 27, OP_ACOS = 28, OP_ATAN = 29, OP_SINH = 30, OP_COSH = 31,
  OP_TANH = 32, OP_ASINH = 33, OP_ACOSH = 34, OP_ATANH = 35, OP_ABS = 36,
  OP_SIGNUM = 37, OP_FLOOR = 38, OP_CEIL = 39, OP_ROUND = 40, OP_TRUNC = 41,
  OP_FRACT = 42, OP_CUSTOM = 43
};
static const uint32_t UNUSED = 0xFFFFFFFFu;

// ---------------------------------------------------------------------------
// Scalar math with Rust std semantics (f32 -> the *f libm variants).
// ---------------------------------------------------------------------------
inline double m_sin(double x) { return std::sin(x); }
inline float m_sin(float x) { return sinf(x); }
inline double m_cos(double x) { return std::cos(x); }
inline float m_cos(float x) { return cosf(x); }
inline double m_tan(double x) { return std::tan(x); }
inline float m_tan(float x) { return tanf(x); }
inline double m_asin(double x) { return std::asin(x); }
inline float m_asin(float x) { return asinf(x); }
inline double m_acos(double x) { return std::acos(x); }
inline float m_acos(float x) { return acosf(x); }
inline double m_atan(double x) { return std::atan(x); }
inline float m_atan(float x) { return atanf(x); }
inline double m_atan2(double a, double b) { return std::atan2(a, b); }
inline float m_atan2(float a, float b) { return atan2f(a, b); }
inline double m_sinh(double x) { return std::sinh(x); }
inline float m_sinh(float x) { return sinhf(x); }
inline double m_cosh(double x) { return std::cosh(x); }
inline float m_cosh(float x) { return coshf(x); }
inline double m_tanh(double x) { return std::tanh(x); }
inline float m_tanh(float x) { return tanhf(x); }
inline double m_exp(double x) { return std::exp(x); }
inline float m_exp(float x) { return expf(x); }
inline double m_exp2(double x) { return std::exp2(x); }
inline float m_exp2(float x) { return exp2f(x); }
inline double m_expm1(double x) { return std::expm1(x); }
inline float m_expm1(float x) { return expm1f(x); }
inline double m_ln(double x) { return std::log(x); }
inline float m_ln(float x) { return logf(x); }
inline double m_log2(double x) { return std::log2(x); }
inline float m_log2(float x) { return log2f(x); }
inline double m_log10(double x) { return std::log10(x); }
inline float m_log10(float x) { return log10f(x); }
inline double m_ln1p(double x) { return std::log1p(x); }
inline float m_ln1p(float x) { return log1pf(x); }
inline double m_sqrt(double x) { return std::sqrt(x); }
inline float m_sqrt(float x) { return sqrtf(x); }
inline double m_cbrt(double x) { return std::cbrt(x); }
inline float m_cbrt(float x) { return cbrtf(x); }
inline double m_hypot(double a, double b) { return std::hypot(a, b); }
inline float m_hypot(float a, float b) { return hypotf(a, b); }
inline double m_pow(double a, double b) { return std::pow(a, b); }
inline float m_pow(float a, float b) { return powf(a, b); }
inline double m_fmod(double a, double b) { return std::fmod(a, b); }
inline float m_fmod(float a, float b) { return fmodf(a, b); }
inline double m_abs(double x) { return std::fabs(x); }
inline float m_abs(float x) { return fabsf(x); }
inline double m_floor(double x) { return std::floor(x); }
inline float m_floor(float x) { return floorf(x); }
inline double m_ceil(double x) { return std::ceil(x); }
inline float m_ceil(float x) { return ceilf(x); }
inline double m_round(double x) { return std::round(x); }  // half away from zero
inline float m_round(float x) { return roundf(x); }
inline double m_trunc(double x) { return std::trunc(x); }
inline float m_trunc(float x) { return truncf(x); }
inline double m_copysign(double a, double b) { return std::copysign(a, b); }
inline float m_copysign(float a, float b) { return copysignf(a, b); }
template <class F> inline bool m_isnan(F x) { return x != x; }
template <class F> inline bool m_isfinite(F x) { return std::isfinite(x); }
// Rust signum: NaN -> NaN, otherwise copysign(1, x) (so -0 -> -1).
template <class F> inline F m_signum(F x) { return m_isnan(x) ? x : m_copysign(F(1), x); }
template <class F> inline F m_fract(F x) { return x - m_trunc(x); }
template <class F> inline F m_recip(F x) { return F(1) / x; }
// Rust std asinh / acosh / atanh identities (not libm).
template <class F> inline F m_asinh(F x) {
  F ax = m_abs(x);
  F ix = F(1) / ax;
  return m_copysign(m_ln1p(ax + (ax / (m_hypot(F(1), ix) + ix))), x);
}
template <class F> inline F m_acosh(F x) {
  if (x < F(1)) return std::numeric_limits<F>::quiet_NaN();
  return m_ln(x + (m_sqrt(x - F(1)) * m_sqrt(x + F(1))));
}
template <class F> inline F m_atanh(F x) {
  return F(0.5) * m_ln1p((F(2) * x) / (F(1) - x));
}
// llvm.powi lowering (compiler-rt __powidf2): repeated squaring.
template <class F> inline F m_powi(F a, int32_t b) {
  const bool recip = b < 0;
  F r = F(1);
  int64_t e = b;  // avoid INT_MIN overflow in the halving
  for (;;) {
    if (e & 1) r *= a;
    e /= 2;
    if (e == 0) break;
    a *= a;
  }
  return recip ? F(1) / r : r;
}
// num-traits to_i32 (truncating, range-checked) followed by the exactness test
// `F::from(ni) == x` used by the integer-exponent fast paths.
template <class F> inline bool m_exact_i32(F x, int32_t* out) {
  if (!(x > F(-2147483649.0) && x < F(2147483648.0))) return false;
  int32_t ni = (int32_t)x;
  if ((F)ni != x) return false;
  *out = ni;
  return true;
}
template <class F> inline F m_nan() { return std::numeric_limits<F>::quiet_NaN(); }
template <class F> inline F m_inf() { return std::numeric_limits<F>::infinity(); }

// ---------------------------------------------------------------------------
// Generic numeric interface: S<T> gives the scalar field, primal access and
// constructors for T in {F, Dual<F>, Taylor<F,K>}.
// ---------------------------------------------------------------------------
template <class F> struct Dual { F re, eps; };
template <class F, int K> struct Taylor { F c[K]; };

template <class T> struct S;
template <> struct S<double> {
  typedef double F;
  static double cst(double v) { return v; }
  static double re(double x) { return x; }
  static bool all_zero(double x) { return x == 0.0; }
};
template <> struct S<float> {
  typedef float F;
  static float cst(double v) { return (float)v; }
  static float re(float x) { return x; }
  static bool all_zero(float x) { return x == 0.0f; }
};
template <class G> struct S<Dual<G>> {
  typedef G F;
  static Dual<G> cst(double v) { return Dual<G>{(G)v, G(0)}; }
  static G re(const Dual<G>& x) { return x.re; }
  static bool all_zero(const Dual<G>& x) { return x.re == G(0) && x.eps == G(0); }
};
template <class G, int K> struct S<Taylor<G, K>> {
  typedef G F;
  static Taylor<G, K> cst(double v) {
    Taylor<G, K> t;
    t.c[0] = (G)v;
    for (int k = 1; k < K; ++k) t.c[k] = G(0);
    return t;
  }
  static G re(const Taylor<G, K>& x) { return x.c[0]; }
  static bool all_zero(const Taylor<G, K>& x) {
    for (int k = 0; k < K; ++k)
      if (x.c[k] != G(0)) return false;
    return true;
  }
};

// ---- Dual<F> arithmetic (src/traits/std_ops.rs:14-83, src/dual.rs) ----
template <class F> inline Dual<F> operator+(Dual<F> a, Dual<F> b) { return {a.re + b.re, a.eps + b.eps}; }
template <class F> inline Dual<F> operator-(Dual<F> a, Dual<F> b) { return {a.re - b.re, a.eps - b.eps}; }
template <class F> inline Dual<F> operator*(Dual<F> a, Dual<F> b) {
  return {a.re * b.re, a.re * b.eps + a.eps * b.re};
}
template <class F> inline Dual<F> operator/(Dual<F> a, Dual<F> b) {
  F inv = F(1) / b.re;
  F re = a.re / b.re;
  return {re, (a.eps - re * b.eps) * inv};
}
template <class F> inline Dual<F> operator-(Dual<F> a) { return {-a.re, -a.eps}; }
template <class F> inline Dual<F> d_rem(Dual<F> a, Dual<F> b) {
  return {m_fmod(a.re, b.re), a.eps - b.eps * m_trunc(a.re / b.re)};
}
template <class F> inline Dual<F> d_const(F v) { return {v, F(0)}; }
template <class F> inline Dual<F> d_chain(Dual<F> x, F f, F df) {
  F eps = (x.eps == F(0)) ? F(0) : x.eps * df;
  return {f, eps};
}

// ---- derivative helpers (src/kernels/mod.rs:85-285), generic in T ----
template <class T> inline void hypot_partials(T a, T b, T r, T& da, T& db);
template <class T> inline void atan2_partials(T a, T b, T& da, T& db);

// ---- per-type elementals used by eval_forward / reverse_partials ----
// Scalar F versions.
#define ORC_SCALAR_ELEM(T)                                                          \
  inline T e_add(T a, T b) { return a + b; }                                        \
  inline T e_sub(T a, T b) { return a - b; }                                        \
  inline T e_mul(T a, T b) { return a * b; }                                        \
  inline T e_div(T a, T b) { return a / b; }                                        \
  inline T e_neg(T a) { return -a; }                                                \
  inline T e_rem(T a, T b) { return m_fmod(a, b); }                                 \
  inline T e_powf(T a, T b) { return m_pow(a, b); }                                 \
  inline T e_atan2(T a, T b) { return m_atan2(a, b); }                              \
  inline T e_hypot(T a, T b) { return m_hypot(a, b); }                              \
  inline T e_recip(T a) { return m_recip(a); }                                      \
  inline T e_sqrt(T a) { return m_sqrt(a); }                                        \
  inline T e_cbrt(T a) { return m_cbrt(a); }                                        \
  inline T e_powi(T a, int32_t n) { return m_powi(a, n); }                          \
  inline T e_exp(T a) { return m_exp(a); }                                          \
  inline T e_exp2(T a) { return m_exp2(a); }                                        \
  inline T e_expm1(T a) { return m_expm1(a); }                                      \
  inline T e_ln(T a) { return m_ln(a); }                                            \
  inline T e_log2(T a) { return m_log2(a); }                                        \
  inline T e_log10(T a) { return m_log10(a); }                                      \
  inline T e_ln1p(T a) { return m_ln1p(a); }                                        \
  inline T e_sin(T a) { return m_sin(a); }                                          \
  inline T e_cos(T a) { return m_cos(a); }                                          \
  inline T e_tan(T a) { return m_tan(a); }                                          \
  inline T e_asin(T a) { return m_asin(a); }                                        \
  inline T e_acos(T a) { return m_acos(a); }                                        \
  inline T e_atan(T a) { return m_atan(a); }                                        \
  inline T e_sinh(T a) { return m_sinh(a); }                                        \
  inline T e_cosh(T a) { return m_cosh(a); }                                        \
  inline T e_tanh(T a) { return m_tanh(a); }                                        \
  inline T e_asinh(T a) { return m_asinh(a); }                                      \
  inline T e_acosh(T a) { return m_acosh(a); }                                      \
  inline T e_atanh(T a) { return m_atanh(a); }                                      \
  inline T e_abs(T a) { return m_abs(a); }                                          \
  inline T e_signum(T a) { return m_signum(a); }                                    \
  inline T e_floor(T a) { return m_floor(a); }                                      \
  inline T e_ceil(T a) { return m_ceil(a); }                                        \
  inline T e_round(T a) { return m_round(a); }                                      \
  inline T e_trunc(T a) { return m_trunc(a); }                                      \
  inline T e_fract(T a) { return m_fract(a); }                                      \
  inline bool e_isnan(T a) { return m_isnan(a); }                                   \
  inline bool e_isfinite(T a) { return m_isfinite(a); }
ORC_SCALAR_ELEM(double)
ORC_SCALAR_ELEM(float)
#undef ORC_SCALAR_ELEM

// Dual<F> elementals (src/dual.rs:44-460).
template <class F> inline Dual<F> e_add(Dual<F> a, Dual<F> b) { return a + b; }
template <class F> inline Dual<F> e_sub(Dual<F> a, Dual<F> b) { return a - b; }
template <class F> inline Dual<F> e_mul(Dual<F> a, Dual<F> b) { return a * b; }
template <class F> inline Dual<F> e_div(Dual<F> a, Dual<F> b) { return a / b; }
template <class F> inline Dual<F> e_neg(Dual<F> a) { return -a; }
template <class F> inline Dual<F> e_rem(Dual<F> a, Dual<F> b) { return d_rem(a, b); }
template <class F> inline Dual<F> e_recip(Dual<F> x) {
  F inv = F(1) / x.re;
  return d_chain(x, inv, -inv * inv);
}
template <class F> inline Dual<F> e_sqrt(Dual<F> x) {
  F s = m_sqrt(x.re);
  F two = F(1) + F(1);
  return d_chain(x, s, F(1) / (two * s));
}
template <class F> inline Dual<F> e_cbrt(Dual<F> x) {
  F c = m_cbrt(x.re);
  F three = F(3);
  return d_chain(x, c, F(1) / (three * c * c));
}
template <class F> inline Dual<F> e_powi(Dual<F> x, int32_t n) {
  if (n == 0) return d_const(F(1));
  F val = m_powi(x.re, n);
  F deriv = (n == INT32_MIN) ? F(n) * val / x.re : F(n) * m_powi(x.re, n - 1);
  return d_chain(x, val, deriv);
}
template <class F> inline F k_ln_deriv(F a) { return (a >= F(0)) ? F(1) / a : m_nan<F>(); }
template <class F> inline F k_log2_deriv(F a) {
  return (a >= F(0)) ? F(1) / (a * (F)0.69314718055994530942) : m_nan<F>();
}
template <class F> inline F k_log10_deriv(F a) {
  return (a >= F(0)) ? F(1) / (a * (F)2.30258509299404568402) : m_nan<F>();
}
template <class F> inline F k_ln1p_deriv(F a) { return (a >= -F(1)) ? F(1) / (F(1) + a) : m_nan<F>(); }
template <class F> inline F k_atanh_deriv(F a) {
  return (a >= -F(1) && a <= F(1)) ? F(1) / ((F(1) - a) * (F(1) + a)) : m_nan<F>();
}
template <class F> inline F k_abs_deriv(F a) { return (a == F(0)) ? F(0) : m_signum(a); }
template <class F> inline F k_atan_deriv(F a) {
  if (m_abs(a) > (F)1e8) {
    F inv = F(1) / a;
    return inv * inv / (F(1) + inv * inv);
  }
  return F(1) / (F(1) + a * a);
}
template <class F> inline F k_asinh_deriv(F a) {
  if (m_abs(a) > (F)1e8) {
    F inv = F(1) / a;
    return m_abs(inv) / m_sqrt(F(1) + inv * inv);
  }
  return F(1) / m_sqrt(a * a + F(1));
}
template <class F> inline F k_acosh_deriv(F a) {
  if (a < F(1)) return m_nan<F>();
  if (a > (F)1e8) {
    F inv = F(1) / a;
    return m_abs(inv) / m_sqrt(F(1) - inv * inv);
  }
  return F(1) / m_sqrt((a - F(1)) * (a + F(1)));
}
template <class F> inline Dual<F> e_exp(Dual<F> x) { F e = m_exp(x.re); return d_chain(x, e, e); }
template <class F> inline Dual<F> e_exp2(Dual<F> x) {
  F e = m_exp2(x.re);
  return d_chain(x, e, e * (F)0.69314718055994530942);
}
template <class F> inline Dual<F> e_expm1(Dual<F> x) { return d_chain(x, m_expm1(x.re), m_exp(x.re)); }
template <class F> inline Dual<F> e_ln(Dual<F> x) { return d_chain(x, m_ln(x.re), k_ln_deriv(x.re)); }
template <class F> inline Dual<F> e_log2(Dual<F> x) { return d_chain(x, m_log2(x.re), k_log2_deriv(x.re)); }
template <class F> inline Dual<F> e_log10(Dual<F> x) { return d_chain(x, m_log10(x.re), k_log10_deriv(x.re)); }
template <class F> inline Dual<F> e_ln1p(Dual<F> x) { return d_chain(x, m_ln1p(x.re), k_ln1p_deriv(x.re)); }
template <class F> inline Dual<F> e_sin(Dual<F> x) { return d_chain(x, m_sin(x.re), m_cos(x.re)); }
template <class F> inline Dual<F> e_cos(Dual<F> x) { return d_chain(x, m_cos(x.re), -m_sin(x.re)); }
template <class F> inline Dual<F> e_tan(Dual<F> x) {
  F c = m_cos(x.re);
  return d_chain(x, m_tan(x.re), F(1) / (c * c));
}
template <class F> inline Dual<F> e_asin(Dual<F> x) {
  return d_chain(x, m_asin(x.re), F(1) / m_sqrt((F(1) - x.re) * (F(1) + x.re)));
}
template <class F> inline Dual<F> e_acos(Dual<F> x) {
  return d_chain(x, m_acos(x.re), -F(1) / m_sqrt((F(1) - x.re) * (F(1) + x.re)));
}
template <class F> inline Dual<F> e_atan(Dual<F> x) { return d_chain(x, m_atan(x.re), k_atan_deriv(x.re)); }
template <class F> inline Dual<F> e_sinh(Dual<F> x) { return d_chain(x, m_sinh(x.re), m_cosh(x.re)); }
template <class F> inline Dual<F> e_cosh(Dual<F> x) { return d_chain(x, m_cosh(x.re), m_sinh(x.re)); }
template <class F> inline Dual<F> e_tanh(Dual<F> x) {
  F c = m_cosh(x.re);
  return d_chain(x, m_tanh(x.re), F(1) / (c * c));
}
template <class F> inline Dual<F> e_asinh(Dual<F> x) { return d_chain(x, m_asinh(x.re), k_asinh_deriv(x.re)); }
template <class F> inline Dual<F> e_acosh(Dual<F> x) { return d_chain(x, m_acosh(x.re), k_acosh_deriv(x.re)); }
template <class F> inline Dual<F> e_atanh(Dual<F> x) { return d_chain(x, m_atanh(x.re), k_atanh_deriv(x.re)); }
template <class F> inline Dual<F> e_abs(Dual<F> x) { return d_chain(x, m_abs(x.re), k_abs_deriv(x.re)); }
template <class F> inline Dual<F> e_signum(Dual<F> x) { return d_const(m_signum(x.re)); }
template <class F> inline Dual<F> e_floor(Dual<F> x) { return d_const(m_floor(x.re)); }
template <class F> inline Dual<F> e_ceil(Dual<F> x) { return d_const(m_ceil(x.re)); }
template <class F> inline Dual<F> e_round(Dual<F> x) { return d_const(m_round(x.re)); }
template <class F> inline Dual<F> e_trunc(Dual<F> x) { return d_const(m_trunc(x.re)); }
template <class F> inline Dual<F> e_fract(Dual<F> x) { return {m_fract(x.re), x.eps}; }
template <class F> inline bool e_isnan(Dual<F> x) { return m_isnan(x.re); }
template <class F> inline bool e_isfinite(Dual<F> x) { return m_isfinite(x.re); }
// Dual::powf (src/dual.rs:358-405)
template <class F> inline Dual<F> e_powf(Dual<F> x, Dual<F> n) {
  if (n.eps == F(0)) {
    int32_t ni;
    if (m_exact_i32(n.re, &ni)) return e_powi(x, ni);
  }
  if (n.re == F(0)) {
    F dy = (x.re > F(0)) ? m_ln(x.re) : F(0);
    return {F(1), dy * n.eps};
  }
  F val = m_pow(x.re, n.re);
  F dx;
  if (x.eps == F(0)) {
    dx = F(0);
  } else if (x.re == F(0) || val == F(0) || !m_isfinite(x.re) || !m_isfinite(val)) {
    dx = n.re * m_pow(x.re, n.re - F(1)) * x.eps;
  } else {
    dx = n.re * val / x.re * x.eps;
  }
  F dy = (n.eps == F(0) || val == F(0) || x.re <= F(0)) ? F(0) : val * m_ln(x.re) * n.eps;
  return {val, dx + dy};
}
// Dual::atan2 (src/dual.rs:418-425) and Dual::hypot (449-460)
template <class F> inline void atan2_partials_s(F a, F b, F& da, F& db) {
  F h = m_hypot(a, b);
  if (h == F(0)) { da = F(0); db = F(0); return; }
  da = b / h / h;
  db = -a / h / h;
}
template <class F> inline Dual<F> e_atan2(Dual<F> x, Dual<F> y) {
  F ds, dot;
  atan2_partials_s(x.re, y.re, ds, dot);
  return {m_atan2(x.re, y.re), ds * x.eps + dot * y.eps};
}
template <class F> inline Dual<F> e_hypot(Dual<F> x, Dual<F> y) {
  F h = m_hypot(x.re, y.re);
  if (h == F(0)) return d_const(h);
  F da = x.re / h, db = y.re / h;
  return {h, da * x.eps + db * y.eps};
}

// ---- Taylor<F,K> elementals (src/taylor.rs, src/taylor_ops.rs) ----
namespace tops {
template <class F, int K> inline void mul(const F* a, const F* b, F* c) {
  for (int k = 0; k < K; ++k) {
    F s = F(0);
    for (int j = 0; j <= k; ++j) s = s + a[j] * b[k - j];
    c[k] = s;
  }
}
template <class F, int K> inline void div(const F* a, const F* b, F* c) {
  c[0] = a[0] / b[0];
  F inv = F(1) / b[0];
  for (int k = 1; k < K; ++k) {
    F s = a[k];
    for (int j = 1; j <= k; ++j) s = s - b[j] * c[k - j];
    c[k] = s * inv;
  }
}
template <class F, int K> inline void recip(const F* a, F* c) {
  F inv = F(1) / a[0];
  c[0] = inv;
  for (int k = 1; k < K; ++k) {
    F s = F(0);
    for (int j = 1; j <= k; ++j) s = s + a[j] * c[k - j];
    c[k] = -s * inv;
  }
}
template <class F, int K> inline void scale(const F* a, F s, F* c) {
  for (int k = 0; k < K; ++k) c[k] = s * a[k];
}
template <class F, int K> inline void exp(const F* a, F* c) {
  c[0] = m_exp(a[0]);
  for (int k = 1; k < K; ++k) {
    F s = F(0);
    for (int j = 1; j <= k; ++j) s = s + F(j) * a[j] * c[k - j];
    c[k] = s / F(k);
  }
}
template <class F, int K> inline void nan_jet(F* c) {
  for (int k = 0; k < K; ++k) c[k] = m_nan<F>();
}
template <class F, int K> inline void vertical(F* c) {
  c[0] = F(0);
  for (int k = 1; k < K; ++k) c[k] = m_inf<F>();
}
template <class F, int K> inline void ln(const F* a, F* c) {
  if (a[0] < F(0)) { nan_jet<F, K>(c); return; }
  F inv = F(1) / a[0];
  c[0] = m_ln(a[0]);
  for (int k = 1; k < K; ++k) {
    F s = F(0);
    for (int j = 1; j < k; ++j) s = s + F(j) * c[j] * a[k - j];
    c[k] = (a[k] - s / F(k)) * inv;
  }
}
template <class F, int K> inline void sqrt(const F* a, F* c) {
  if (a[0] == F(0)) { vertical<F, K>(c); return; }
  if (a[0] < F(0)) { nan_jet<F, K>(c); return; }
  c[0] = m_sqrt(a[0]);
  F two_c0 = F(2) * c[0];
  for (int k = 1; k < K; ++k) {
    F s = F(0);
    for (int j = 1; j < k; ++j) s = s + c[j] * c[k - j];
    c[k] = (a[k] - s) / two_c0;
  }
}
template <class F, int K> inline void sin_cos(const F* a, F* s, F* co) {
  s[0] = m_sin(a[0]);
  co[0] = m_cos(a[0]);
  for (int k = 1; k < K; ++k) {
    F inv_k = F(1) / F(k);
    F ss = F(0), sc = F(0);
    for (int j = 1; j <= k; ++j) {
      F jf = F(j);
      ss = ss + jf * a[j] * co[k - j];
      sc = sc + jf * a[j] * s[k - j];
    }
    s[k] = ss * inv_k;
    co[k] = -sc * inv_k;
  }
}
template <class F, int K> inline void sinh_cosh(const F* a, F* sh, F* ch) {
  sh[0] = m_sinh(a[0]);
  ch[0] = m_cosh(a[0]);
  for (int k = 1; k < K; ++k) {
    F inv_k = F(1) / F(k);
    F ss = F(0), sc = F(0);
    for (int j = 1; j <= k; ++j) {
      F jf = F(j);
      ss = ss + jf * a[j] * ch[k - j];
      sc = sc + jf * a[j] * sh[k - j];
    }
    sh[k] = ss * inv_k;
    ch[k] = sc * inv_k;
  }
}
template <class F, int K> inline void integrate(const F* a, const F* g, F* c) {
  for (int k = 1; k < K; ++k) {
    F s = F(0);
    for (int j = 1; j <= k; ++j) s = s + F(j) * a[j] * g[k - j];
    c[k] = s / F(k);
  }
}
template <class F, int K> inline void atan(const F* a, F* c) {
  F s1[K], s2[K];
  mul<F, K>(a, a, s1);
  for (int k = 0; k < K; ++k) s2[k] = s1[k];
  s2[0] = F(1) + s1[0];
  c[0] = m_atan(a[0]);
  recip<F, K>(s2, s1);
  integrate<F, K>(a, s1, c);
}
template <class F, int K> inline void asin(const F* a, F* c) {
  F s1[K], s2[K];
  c[0] = m_asin(a[0]);
  mul<F, K>(a, a, s1);
  s2[0] = (F(1) - a[0]) * (F(1) + a[0]);
  for (int k = 1; k < K; ++k) s2[k] = -s1[k];
  sqrt<F, K>(s2, s1);
  recip<F, K>(s1, s2);
  integrate<F, K>(a, s2, c);
}
template <class F, int K> inline void acos(const F* a, F* c) {
  asin<F, K>(a, c);
  c[0] = m_acos(a[0]);
  for (int k = 1; k < K; ++k) c[k] = -c[k];
}
template <class F, int K> inline void tan(const F* a, F* c) {
  F sc[K];
  c[0] = m_tan(a[0]);
  sc[0] = F(1) + c[0] * c[0];
  for (int k = 1; k < K; ++k) {
    F s = F(0);
    for (int j = 1; j <= k; ++j) s = s + F(j) * a[j] * sc[k - j];
    c[k] = s / F(k);
    F sk = F(0);
    for (int j = 0; j <= k; ++j) sk = sk + c[j] * c[k - j];
    sc[k] = sk;
  }
}
template <class F, int K> inline void tanh(const F* a, F* c) {
  F sc[K];
  c[0] = m_tanh(a[0]);
  sc[0] = F(1) - c[0] * c[0];
  for (int k = 1; k < K; ++k) {
    F s = F(0);
    for (int j = 1; j <= k; ++j) s = s + F(j) * a[j] * sc[k - j];
    c[k] = s / F(k);
    F sk = F(0);
    for (int j = 0; j <= k; ++j) sk = sk + c[j] * c[k - j];
    sc[k] = -sk;
  }
}
template <class F, int K> inline void asinh(const F* a, F* c) {
  F s1[K], s2[K];
  c[0] = m_asinh(a[0]);
  mul<F, K>(a, a, s1);
  for (int k = 0; k < K; ++k) s2[k] = s1[k];
  s2[0] = F(1) + s1[0];
  sqrt<F, K>(s2, s1);
  recip<F, K>(s1, s2);
  integrate<F, K>(a, s2, c);
}
template <class F, int K> inline void acosh(const F* a, F* c) {
  if (a[0] < F(1)) { nan_jet<F, K>(c); return; }
  F s1[K], s2[K];
  c[0] = m_acosh(a[0]);
  mul<F, K>(a, a, s1);
  for (int k = 0; k < K; ++k) s2[k] = s1[k];
  s2[0] = (a[0] - F(1)) * (a[0] + F(1));
  sqrt<F, K>(s2, s1);
  recip<F, K>(s1, s2);
  integrate<F, K>(a, s2, c);
}
template <class F, int K> inline void atanh(const F* a, F* c) {
  if (a[0] < -F(1) || a[0] > F(1)) { nan_jet<F, K>(c); return; }
  F s1[K], s2[K];
  c[0] = m_atanh(a[0]);
  mul<F, K>(a, a, s1);
  s2[0] = (F(1) - a[0]) * (F(1) + a[0]);
  for (int k = 1; k < K; ++k) s2[k] = -s1[k];
  recip<F, K>(s2, s1);
  integrate<F, K>(a, s1, c);
}
template <class F, int K> inline void powi_squaring(const F* a, int32_t n, F* c) {
  F s1[K], s2[K];
  uint32_t abs_n = (n < 0) ? (uint32_t)(-(int64_t)n) : (uint32_t)n;
  c[0] = F(1);
  for (int k = 1; k < K; ++k) c[k] = F(0);
  for (int k = 0; k < K; ++k) s1[k] = a[k];
  uint32_t power = abs_n;
  while (power > 0) {
    if (power & 1u) {
      mul<F, K>(c, s1, s2);
      for (int k = 0; k < K; ++k) c[k] = s2[k];
    }
    power >>= 1;
    if (power > 0) {
      for (int k = 0; k < K; ++k) {
        F s = F(0);
        for (int j = 0; j <= k; ++j) s = s + s1[j] * s1[k - j];
        s2[k] = s;
      }
      for (int k = 0; k < K; ++k) s1[k] = s2[k];
    }
  }
  if (n < 0) {
    for (int k = 0; k < K; ++k) s1[k] = c[k];
    recip<F, K>(s1, c);
  }
}
template <class F, int K> inline void powi(const F* a, int32_t n, F* c) {
  if (n == 0) {
    c[0] = F(1);
    for (int k = 1; k < K; ++k) c[k] = F(0);
    return;
  }
  if (n == 1) { for (int k = 0; k < K; ++k) c[k] = a[k]; return; }
  if (n == -1) { recip<F, K>(a, c); return; }
  uint32_t abs_n = (n < 0) ? (uint32_t)(-(int64_t)n) : (uint32_t)n;
  if (a[0] <= F(0) || abs_n <= 8u) {
    powi_squaring<F, K>(a, n, c);
  } else {
    F s1[K], s2[K];
    ln<F, K>(a, s1);
    scale<F, K>(s1, F(n), s2);
    exp<F, K>(s2, c);
    c[0] = m_powi(a[0], n);
  }
}
template <class F, int K> inline void powf(const F* a, const F* b, F* c) {
  bool const_b = true;
  for (int k = 1; k < K; ++k)
    if (b[k] != F(0)) const_b = false;
  int32_t ni;
  if (const_b && m_exact_i32(b[0], &ni)) { powi<F, K>(a, ni, c); return; }
  if (a[0] == F(0)) {
    F b0 = b[0];
    if (m_exact_i32(b0, &ni)) { nan_jet<F, K>(c); return; }
    c[0] = m_pow(a[0], b0);
    for (int k = 1; k < K; ++k) c[k] = (F(k) < b0) ? F(0) : m_inf<F>();
    return;
  }
  F s1[K], s2[K];
  ln<F, K>(a, s1);
  mul<F, K>(b, s1, s2);
  exp<F, K>(s2, c);
  if (a[0] < F(0)) c[0] = m_nan<F>();
  else c[0] = m_pow(a[0], b[0]);
}
template <class F, int K> inline void cbrt(const F* a, F* c) {
  if (a[0] == F(0)) { vertical<F, K>(c); return; }
  F s1[K], s2[K];
  F third = F(1) / F(3);
  if (a[0] < F(0)) {
    for (int i = 0; i < K; ++i) c[i] = -a[i];
    ln<F, K>(c, s1);
    scale<F, K>(s1, third, s2);
    exp<F, K>(s2, c);
    c[0] = m_cbrt(a[0]);
    for (int i = 1; i < K; ++i) c[i] = -c[i];
  } else {
    ln<F, K>(a, s1);
    scale<F, K>(s1, third, s2);
    exp<F, K>(s2, c);
    c[0] = m_cbrt(a[0]);
  }
}
template <class F, int K> inline void exp2(const F* a, F* c) {
  F s[K];
  F ln2 = m_ln(F(2));
  scale<F, K>(a, ln2, s);
  exp<F, K>(s, c);
  c[0] = m_exp2(a[0]);
}
template <class F, int K> inline void expm1(const F* a, F* c) {
  exp<F, K>(a, c);
  c[0] = m_expm1(a[0]);
}
template <class F, int K> inline void log2(const F* a, F* c) {
  ln<F, K>(a, c);
  F inv = F(1) / m_ln(F(2));
  c[0] = m_log2(a[0]);
  for (int k = 1; k < K; ++k) c[k] = c[k] * inv;
}
template <class F, int K> inline void log10(const F* a, F* c) {
  ln<F, K>(a, c);
  F inv = F(1) / m_ln(F(10));
  c[0] = m_log10(a[0]);
  for (int k = 1; k < K; ++k) c[k] = c[k] * inv;
}
template <class F, int K> inline void ln1p(const F* a, F* c) {
  F s[K];
  for (int k = 1; k < K; ++k) s[k] = a[k];
  s[0] = F(1) + a[0];
  ln<F, K>(s, c);
  c[0] = m_ln1p(a[0]);
}
template <class F, int K> inline void hypot(const F* a, const F* b, F* c) {
  int m = 0;
  F amax = m_abs(a[0]), bmax = m_abs(b[0]);
  // IEEE maxNum: NaN operands are dropped.
  F mx = m_isnan(amax) ? bmax : (m_isnan(bmax) ? amax : (amax > bmax ? amax : bmax));
  if (mx == F(0)) {
    if (m_isnan(a[0]) || m_isnan(b[0])) { nan_jet<F, K>(c); return; }
    int found = -1;
    for (int k = 1; k < K; ++k)
      if (a[k] != F(0) || b[k] != F(0)) { found = k; break; }
    if (found < 0) { for (int k = 0; k < K; ++k) c[k] = F(0); return; }
    m = found;
    if (m_isnan(a[m]) || m_isnan(b[m])) {
      for (int k = 0; k < m; ++k) c[k] = F(0);
      c[m] = m_hypot(a[m], b[m]);
      for (int k = m + 1; k < K; ++k) c[k] = m_nan<F>();
      return;
    }
  }
  F am = m_abs(a[m]), bm = m_abs(b[m]);
  F scl = m_isnan(am) ? bm : (m_isnan(bm) ? am : (am > bm ? am : bm));
  F inv = F(1) / scl;
  F s1[K], s2[K];
  for (int k = 0; k < K; ++k) s1[k] = (k + m < K) ? a[k + m] * inv : F(0);
  for (int k = 0; k < K; ++k) s2[k] = (k + m < K) ? b[k + m] * inv : F(0);
  mul<F, K>(s1, s1, c);
  mul<F, K>(s2, s2, s1);
  for (int k = 0; k < K; ++k) c[k] = c[k] + s1[k];
  sqrt<F, K>(c, s1);
  for (int k = 0; k < K; ++k) c[k] = s1[k] * scl;
  c[0] = m_hypot(a[m], b[m]);
  if (m > 0) {
    for (int k = K - 1; k >= m; --k) c[k] = c[k - m];
    for (int k = 0; k < m; ++k) c[k] = F(0);
  }
}
template <class F, int K> inline void atan2(const F* a, const F* b, F* c) {
  F s1[K];
  if (b[0] != F(0)) {
    div<F, K>(a, b, s1);
    atan<F, K>(s1, c);
    c[0] = m_atan2(a[0], b[0]);
  } else if (a[0] != F(0)) {
    div<F, K>(b, a, s1);
    atan<F, K>(s1, c);
    for (int k = 0; k < K; ++k) c[k] = -c[k];
    c[0] = m_atan2(a[0], b[0]);
  } else {
    c[0] = F(0);
    for (int k = 1; k < K; ++k) c[k] = F(0);
  }
}
}  // namespace tops

#define ORC_TAY_UN(name, ...)                                              \
  template <class F, int K> inline Taylor<F, K> name(Taylor<F, K> x) {    \
    Taylor<F, K> r;                                                       \
    __VA_ARGS__;                                                          \
    return r;                                                             \
  }
template <class F, int K> inline Taylor<F, K> e_add(Taylor<F, K> a, Taylor<F, K> b) {
  Taylor<F, K> r; for (int k = 0; k < K; ++k) r.c[k] = a.c[k] + b.c[k]; return r;
}
template <class F, int K> inline Taylor<F, K> e_sub(Taylor<F, K> a, Taylor<F, K> b) {
  Taylor<F, K> r; for (int k = 0; k < K; ++k) r.c[k] = a.c[k] - b.c[k]; return r;
}
template <class F, int K> inline Taylor<F, K> e_mul(Taylor<F, K> a, Taylor<F, K> b) {
  Taylor<F, K> r; tops::mul<F, K>(a.c, b.c, r.c); return r;
}
template <class F, int K> inline Taylor<F, K> e_div(Taylor<F, K> a, Taylor<F, K> b) {
  Taylor<F, K> r; tops::div<F, K>(a.c, b.c, r.c); return r;
}
template <class F, int K> inline Taylor<F, K> e_neg(Taylor<F, K> a) {
  Taylor<F, K> r; for (int k = 0; k < K; ++k) r.c[k] = -a.c[k]; return r;
}
template <class F, int K> inline Taylor<F, K> e_rem(Taylor<F, K> a, Taylor<F, K> b) {
  Taylor<F, K> r;
  if (b.c[0] == F(0)) { tops::nan_jet<F, K>(r.c); return r; }
  F q = m_trunc(a.c[0] / b.c[0]);
  r.c[0] = m_fmod(a.c[0], b.c[0]);
  for (int k = 1; k < K; ++k) r.c[k] = a.c[k] - b.c[k] * q;
  return r;
}
template <class F, int K> inline Taylor<F, K> e_powf(Taylor<F, K> a, Taylor<F, K> b) {
  Taylor<F, K> r; tops::powf<F, K>(a.c, b.c, r.c); return r;
}
template <class F, int K> inline Taylor<F, K> e_atan2(Taylor<F, K> a, Taylor<F, K> b) {
  Taylor<F, K> r; tops::atan2<F, K>(a.c, b.c, r.c); return r;
}
template <class F, int K> inline Taylor<F, K> e_hypot(Taylor<F, K> a, Taylor<F, K> b) {
  Taylor<F, K> r; tops::hypot<F, K>(a.c, b.c, r.c); return r;
}
template <class F, int K> inline Taylor<F, K> e_powi(Taylor<F, K> a, int32_t n) {
  Taylor<F, K> r; tops::powi<F, K>(a.c, n, r.c); return r;
}
ORC_TAY_UN(e_recip, tops::recip<F, K>(x.c, r.c))
ORC_TAY_UN(e_sqrt, tops::sqrt<F, K>(x.c, r.c))
ORC_TAY_UN(e_cbrt, tops::cbrt<F, K>(x.c, r.c))
ORC_TAY_UN(e_exp, tops::exp<F, K>(x.c, r.c))
ORC_TAY_UN(e_exp2, tops::exp2<F, K>(x.c, r.c))
ORC_TAY_UN(e_expm1, tops::expm1<F, K>(x.c, r.c))
ORC_TAY_UN(e_ln, tops::ln<F, K>(x.c, r.c))
ORC_TAY_UN(e_log2, tops::log2<F, K>(x.c, r.c))
ORC_TAY_UN(e_log10, tops::log10<F, K>(x.c, r.c))
ORC_TAY_UN(e_ln1p, tops::ln1p<F, K>(x.c, r.c))
ORC_TAY_UN(e_tan, tops::tan<F, K>(x.c, r.c))
ORC_TAY_UN(e_asin, tops::asin<F, K>(x.c, r.c))
ORC_TAY_UN(e_acos, tops::acos<F, K>(x.c, r.c))
ORC_TAY_UN(e_atan, tops::atan<F, K>(x.c, r.c))
ORC_TAY_UN(e_tanh, tops::tanh<F, K>(x.c, r.c))
ORC_TAY_UN(e_asinh, tops::asinh<F, K>(x.c, r.c))
ORC_TAY_UN(e_acosh, tops::acosh<F, K>(x.c, r.c))
ORC_TAY_UN(e_atanh, tops::atanh<F, K>(x.c, r.c))
#undef ORC_TAY_UN
template <class F, int K> inline Taylor<F, K> e_sin(Taylor<F, K> x) {
  Taylor<F, K> s, c; tops::sin_cos<F, K>(x.c, s.c, c.c); return s;
}
template <class F, int K> inline Taylor<F, K> e_cos(Taylor<F, K> x) {
  Taylor<F, K> s, c; tops::sin_cos<F, K>(x.c, s.c, c.c); return c;
}
template <class F, int K> inline Taylor<F, K> e_sinh(Taylor<F, K> x) {
  Taylor<F, K> s, c; tops::sinh_cosh<F, K>(x.c, s.c, c.c); return s;
}
template <class F, int K> inline Taylor<F, K> e_cosh(Taylor<F, K> x) {
  Taylor<F, K> s, c; tops::sinh_cosh<F, K>(x.c, s.c, c.c); return c;
}
template <class F, int K> inline Taylor<F, K> e_abs(Taylor<F, K> x) {
  F sign = F(1);
  if (x.c[0] != F(0)) sign = m_signum(x.c[0]);
  else {
    for (int k = 1; k < K; ++k)
      if (x.c[k] != F(0)) { sign = m_signum(x.c[k]); break; }
  }
  Taylor<F, K> r;
  for (int k = 0; k < K; ++k) r.c[k] = x.c[k] * sign;
  return r;
}
template <class F, int K> inline Taylor<F, K> tay_const(F v) {
  Taylor<F, K> r; r.c[0] = v; for (int k = 1; k < K; ++k) r.c[k] = F(0); return r;
}
template <class F, int K> inline Taylor<F, K> e_signum(Taylor<F, K> x) { return tay_const<F, K>(m_signum(x.c[0])); }
template <class F, int K> inline Taylor<F, K> e_floor(Taylor<F, K> x) { return tay_const<F, K>(m_floor(x.c[0])); }
template <class F, int K> inline Taylor<F, K> e_ceil(Taylor<F, K> x) { return tay_const<F, K>(m_ceil(x.c[0])); }
template <class F, int K> inline Taylor<F, K> e_round(Taylor<F, K> x) { return tay_const<F, K>(m_round(x.c[0])); }
template <class F, int K> inline Taylor<F, K> e_trunc(Taylor<F, K> x) { return tay_const<F, K>(m_trunc(x.c[0])); }
template <class F, int K> inline Taylor<F, K> e_fract(Taylor<F, K> x) { Taylor<F, K> r = x; r.c[0] = m_fract(x.c[0]); return r; }
template <class F, int K> inline bool e_isnan(Taylor<F, K> x) { return m_isnan(x.c[0]); }
template <class F, int K> inline bool e_isfinite(Taylor<F, K> x) { return m_isfinite(x.c[0]); }

// ---------------------------------------------------------------------------
// eval_forward (src/opcode.rs:228-318), generic in T.
// ---------------------------------------------------------------------------
template <class T> inline T eval_forward(uint32_t op, T a, T b) {
  typedef S<T> St;
  switch (op) {
    case OP_ADD: return e_add(a, b);
    case OP_SUB: return e_sub(a, b);
    case OP_MUL: return e_mul(a, b);
    case OP_DIV: return e_div(a, b);
    case OP_REM: return e_rem(a, b);
    case OP_POWF: return e_powf(a, b);
    case OP_ATAN2: return e_atan2(a, b);
    case OP_HYPOT: return e_hypot(a, b);
    case OP_MAX: return (St::re(a) >= St::re(b) || e_isnan(b)) ? a : b;
    case OP_MIN: return (St::re(a) <= St::re(b) || e_isnan(b)) ? a : b;
    case OP_NEG: return e_neg(a);
    case OP_RECIP: return e_recip(a);
    case OP_SQRT: return e_sqrt(a);
    case OP_CBRT: return e_cbrt(a);
    case OP_EXP: return e_exp(a);
    case OP_EXP2: return e_exp2(a);
    case OP_EXPM1: return e_expm1(a);
    case OP_LN: return e_ln(a);
    case OP_LOG2: return e_log2(a);
    case OP_LOG10: return e_log10(a);
    case OP_LN1P: return e_ln1p(a);
    case OP_SIN: return e_sin(a);
    case OP_COS: return e_cos(a);
    case OP_TAN: return e_tan(a);
    case OP_ASIN: return e_asin(a);
    case OP_ACOS: return e_acos(a);
    case OP_ATAN: return e_atan(a);
    case OP_SINH: return e_sinh(a);
    case OP_COSH: return e_cosh(a);
    case OP_TANH: return e_tanh(a);
    case OP_ASINH: return e_asinh(a);
    case OP_ACOSH: return e_acosh(a);
    case OP_ATANH: return e_atanh(a);
    case OP_ABS: return e_abs(a);
    case OP_SIGNUM: return e_signum(a);
    case OP_FLOOR: return e_floor(a);
    case OP_CEIL: return e_ceil(a);
    case OP_ROUND: return e_round(a);
    case OP_TRUNC: return e_trunc(a);
    case OP_FRACT: return e_fract(a);
    default: return St::cst(std::numeric_limits<double>::quiet_NaN());
  }
}

// Generic helpers for reverse_partials over T in {F, Dual<F>}.
template <class F> inline F g_nan(F) { return m_nan<F>(); }
template <class F> inline Dual<F> g_nan(Dual<F>) { return d_const(m_nan<F>()); }

// powi_reverse_partial (src/opcode.rs:503-511)
template <class T> inline T powi_reverse_partial(int32_t e, T a, T r) {
  typedef S<T> St;
  if (e == 0) return St::cst(0.0);
  if (e == INT32_MIN) return e_div(e_mul(St::cst((double)e), r), a);
  return e_mul(St::cst((double)e), e_powi(a, e - 1));
}

// reverse_partials (src/opcode.rs:331-483) generic in T.
template <class T> inline void reverse_partials(uint32_t op, T a, T b, T r, T& da, T& db) {
  typedef S<T> St;
  const T zero = St::cst(0.0), one = St::cst(1.0);
  da = zero;
  db = zero;
  switch (op) {
    case OP_ADD: da = one; db = one; return;
    case OP_SUB: da = one; db = e_neg(one); return;
    case OP_MUL: da = b; db = a; return;
    case OP_DIV: {
      T inv = e_div(one, b);
      da = inv;
      db = e_mul(e_neg(r), inv);
      return;
    }
    case OP_REM: da = one; db = e_neg(e_trunc(e_div(a, b))); return;
    case OP_POWF: {
      if (St::re(b) == 0) {
        db = (St::re(a) > 0) ? e_ln(a) : zero;
        da = zero;
        return;
      }
      if (St::re(a) == 0 || St::re(r) == 0 || !e_isfinite(a) || !e_isfinite(r))
        da = e_mul(b, e_powf(a, e_sub(b, one)));
      else
        da = e_div(e_mul(b, r), a);
      db = (St::re(r) == 0 || St::re(a) <= 0) ? zero : e_mul(r, e_ln(a));
      return;
    }
    case OP_ATAN2: {
      T h = e_hypot(a, b);
      if (St::re(h) == 0) return;
      da = e_div(e_div(b, h), h);
      db = e_div(e_div(e_neg(a), h), h);
      return;
    }
    case OP_HYPOT: {
      if (St::re(r) == 0) return;
      da = e_div(a, r);
      db = e_div(b, r);
      return;
    }
    case OP_MAX:
      if (St::re(a) >= St::re(b) || e_isnan(b)) da = one; else db = one;
      return;
    case OP_MIN:
      if (St::re(a) <= St::re(b) || e_isnan(b)) da = one; else db = one;
      return;
    case OP_NEG: da = e_neg(one); return;
    case OP_RECIP: {
      T inv = e_div(one, a);
      da = e_mul(e_neg(inv), inv);
      return;
    }
    case OP_SQRT: {
      T two = e_add(one, one);
      da = e_div(one, e_mul(two, r));
      return;
    }
    case OP_CBRT: {
      T three = St::cst(3.0);
      da = e_div(one, e_mul(e_mul(three, r), r));
      return;
    }
    case OP_EXP: da = r; return;
    case OP_EXP2: da = e_mul(r, St::cst(0.69314718055994530942)); return;
    case OP_EXPM1: da = e_add(r, one); return;
    case OP_LN: da = (St::re(a) >= 0) ? e_div(one, a) : g_nan(a); return;
    case OP_LOG2:
      da = (St::re(a) >= 0) ? e_div(one, e_mul(a, St::cst(0.69314718055994530942))) : g_nan(a);
      return;
    case OP_LOG10:
      da = (St::re(a) >= 0) ? e_div(one, e_mul(a, St::cst(2.30258509299404568402))) : g_nan(a);
      return;
    case OP_LN1P: da = (St::re(a) >= -1) ? e_div(one, e_add(one, a)) : g_nan(a); return;
    case OP_SIN: da = e_cos(a); return;
    case OP_COS: da = e_neg(e_sin(a)); return;
    case OP_TAN: {
      T c = e_cos(a);
      da = e_div(one, e_mul(c, c));
      return;
    }
    case OP_ASIN: da = e_div(one, e_sqrt(e_mul(e_sub(one, a), e_add(one, a)))); return;
    case OP_ACOS: da = e_div(e_neg(one), e_sqrt(e_mul(e_sub(one, a), e_add(one, a)))); return;
    case OP_ATAN: {
      if (m_abs(St::re(a)) > (typename St::F)1e8) {
        T inv = e_div(one, a);
        da = e_div(e_mul(inv, inv), e_add(one, e_mul(inv, inv)));
      } else {
        da = e_div(one, e_add(one, e_mul(a, a)));
      }
      return;
    }
    case OP_SINH: da = e_cosh(a); return;
    case OP_COSH: da = e_sinh(a); return;
    case OP_TANH: {
      T c = e_cosh(a);
      da = e_div(one, e_mul(c, c));
      return;
    }
    case OP_ASINH: {
      if (m_abs(St::re(a)) > (typename St::F)1e8) {
        T inv = e_div(one, a);
        da = e_div(e_abs(inv), e_sqrt(e_add(one, e_mul(inv, inv))));
      } else {
        da = e_div(one, e_sqrt(e_add(e_mul(a, a), one)));
      }
      return;
    }
    case OP_ACOSH: {
      if (St::re(a) < 1) { da = g_nan(a); return; }
      if (St::re(a) > (typename St::F)1e8) {
        T inv = e_div(one, a);
        da = e_div(e_abs(inv), e_sqrt(e_sub(one, e_mul(inv, inv))));
      } else {
        da = e_div(one, e_sqrt(e_mul(e_sub(a, one), e_add(a, one))));
      }
      return;
    }
    case OP_ATANH:
      if (St::re(a) >= -1 && St::re(a) <= 1)
        da = e_div(one, e_mul(e_sub(one, a), e_add(one, a)));
      else
        da = g_nan(a);
      return;
    case OP_ABS: da = (St::re(a) == 0) ? zero : e_signum(a); return;
    case OP_SIGNUM: case OP_FLOOR: case OP_CEIL: case OP_ROUND: case OP_TRUNC: return;
    case OP_FRACT: da = one; return;
    default: return;
  }
}

// ---------------------------------------------------------------------------
// Flattened tape view (same arrays the GPU upload receives).
// ---------------------------------------------------------------------------
struct TapeView {
  const uint32_t* opcodes;
  const uint32_t* arg0;
  const uint32_t* arg1;
  const double* values;  // recorded values (constants); f64
  uint32_t num_ops;
  uint32_t num_inputs;
  const uint32_t* output_indices;
  uint32_t num_outputs;
};

// forward_into + forward_dispatch (src/bytecode_tape/forward.rs:11-43, 194-229)
template <class F> inline void forward_into(const TapeView& t, const F* x, std::vector<F>& v) {
  v.resize(t.num_ops);
  for (uint32_t i = 0; i < t.num_ops; ++i) v[i] = (F)t.values[i];
  for (uint32_t i = 0; i < t.num_inputs; ++i) v[i] = x[i];
  for (uint32_t i = 0; i < t.num_ops; ++i) {
    uint32_t op = t.opcodes[i];
    if (op == OP_INPUT || op == OP_CONST) continue;
    uint32_t ai = t.arg0[i], bi = t.arg1[i];
    F a = v[ai];
    if (op == OP_POWI) { v[i] = m_powi(a, (int32_t)bi); continue; }
    F b = (bi != UNUSED) ? v[bi] : F(0);
    v[i] = eval_forward<F>(op, a, b);
  }
}

// reverse_sweep_core (src/bytecode_tape/reverse.rs:32-105) seeded at `seed`.
template <class F> inline void reverse_from(const TapeView& t, const std::vector<F>& v, uint32_t seed,
                                            std::vector<F>& adj) {
  adj.assign(t.num_ops, F(0));
  adj[seed] = F(1);
  for (int64_t ii = (int64_t)t.num_ops - 1; ii >= 0; --ii) {
    uint32_t i = (uint32_t)ii;
    F g = adj[i];
    if (g == F(0)) continue;
    uint32_t op = t.opcodes[i];
    if (op == OP_INPUT || op == OP_CONST) continue;
    adj[i] = F(0);
    uint32_t ai = t.arg0[i], bi = t.arg1[i];
    F a = v[ai];
    if (op == OP_POWI) {
      F da = powi_reverse_partial<F>((int32_t)bi, a, v[i]);
      if (da != F(0)) adj[ai] = adj[ai] + da * g;
      continue;
    }
    F b = (bi != UNUSED) ? v[bi] : F(0);
    F r = v[i];
    F da, db;
    reverse_partials<F>(op, a, b, r, da, db);
    if (da != F(0)) adj[ai] = adj[ai] + da * g;
    if (bi != UNUSED && db != F(0)) adj[bi] = adj[bi] + db * g;
  }
}

// forward_tangent_inner (src/bytecode_tape/tangent.rs:97-147) for T = Dual or Taylor.
template <class T> inline void forward_tangent(const TapeView& t, const T* in, std::vector<T>& buf) {
  typedef S<T> St;
  buf.resize(t.num_ops);
  uint32_t k = 0;
  for (uint32_t i = 0; i < t.num_ops; ++i) {
    uint32_t op = t.opcodes[i];
    if (op == OP_INPUT) { buf[i] = in[k++]; continue; }
    if (op == OP_CONST) { buf[i] = St::cst(t.values[i]); continue; }
    uint32_t ai = t.arg0[i], bi = t.arg1[i];
    T a = buf[ai];
    if (op == OP_POWI) { buf[i] = e_powi(a, (int32_t)bi); continue; }
    T b = (bi != UNUSED) ? buf[bi] : St::cst(0.0);
    buf[i] = eval_forward<T>(op, a, b);
  }
}

// reverse_tangent_inner (src/bytecode_tape/tangent.rs:191-262) for T = Dual<F>.
template <class F> inline void reverse_tangent(const TapeView& t, const std::vector<Dual<F>>& tv,
                                               std::vector<Dual<F>>& buf) {
  typedef Dual<F> T;
  typedef S<T> St;
  buf.assign(t.num_ops, St::cst(0.0));
  uint32_t seed = t.output_indices ? t.output_indices[0] : 0;
  buf[seed] = St::cst(1.0);
  for (int64_t ii = (int64_t)t.num_ops - 1; ii >= 0; --ii) {
    uint32_t i = (uint32_t)ii;
    uint32_t op = t.opcodes[i];
    if (op == OP_INPUT || op == OP_CONST) continue;
    T g = buf[i];
    if (St::all_zero(g)) continue;
    buf[i] = St::cst(0.0);
    uint32_t ai = t.arg0[i], bi = t.arg1[i];
    T a = tv[ai];
    if (op == OP_POWI) {
      T da = powi_reverse_partial<T>((int32_t)bi, a, tv[i]);
      if (!St::all_zero(da)) buf[ai] = buf[ai] + da * g;
      continue;
    }
    T b = (bi != UNUSED) ? tv[bi] : St::cst(0.0);
    T r = tv[i];
    T da, db;
    reverse_partials<T>(op, a, b, r, da, db);
    if (!St::all_zero(da)) buf[ai] = buf[ai] + da * g;
    if (bi != UNUSED && !St::all_zero(db)) buf[bi] = buf[bi] + db * g;
  }
}

// hvp (src/bytecode_tape/tangent.rs:268-310)
template <class F> inline void hvp(const TapeView& t, const F* x, const F* v, F* grad, F* hv) {
  std::vector<Dual<F>> in(t.num_inputs), tv, adj;
  for (uint32_t i = 0; i < t.num_inputs; ++i) in[i] = Dual<F>{x[i], v[i]};
  forward_tangent<Dual<F>>(t, in.data(), tv);
  reverse_tangent<F>(t, tv, adj);
  for (uint32_t i = 0; i < t.num_inputs; ++i) {
    grad[i] = adj[i].re;
    hv[i] = adj[i].eps;
  }
}

// Magnitudes of the terms behind each result (no reference counterpart;
// the scale a parity bar is stated against where results cancel).
// Forward, every entry r = f(a, b) carries
//   E_r = |da| E_a + |db| E_b + |r|    (E = |x_i| at inputs, |c| at constants)
//   T_r = |da| T_a + |db| T_b          (T = |v_i| at inputs, 0 at constants)
// with da, db the reverse partials: E / 2^53 bounds the rounding error of
// the value to first order (running error analysis) and T is the sum of the
// absolute terms of the tangent grad . v.  Reverse, the adjoints of the
// absolute partials (evaluated as Dual numbers on (value, T), each component
// taken in absolute value) give, per input, the sum of the absolute terms of
// the gradient (.re) and of H v (.eps); c2 = v^T H v / 2 gets
// sum_i |v_i| (H v)_i^abs / 2.
template <class F> inline void magnitudes(const TapeView& t, const F* x, const F* v, F* val_mag /*[m]*/,
                                          F* grad_mag /*[n]*/, F* hv_mag /*[n]*/, F* c1_mag, F* c2_mag) {
  typedef Dual<F> D;
  std::vector<F> val;
  forward_into<F>(t, x, val);
  const uint32_t N = t.num_ops;
  std::vector<F> E(N), T(N);
  for (uint32_t i = 0; i < N; ++i) {
    uint32_t op = t.opcodes[i];
    if (op == OP_INPUT) { E[i] = m_abs(x[i]); T[i] = m_abs(v[i]); continue; }
    if (op == OP_CONST) { E[i] = m_abs((F)t.values[i]); T[i] = F(0); continue; }
    uint32_t ai = t.arg0[i], bi = t.arg1[i];
    F a = val[ai], r = val[i];
    F da, db = F(0);
    if (op == OP_POWI) {
      da = powi_reverse_partial<F>((int32_t)bi, a, r);
      E[i] = m_abs(da) * E[ai] + m_abs(r);
      T[i] = m_abs(da) * T[ai];
      continue;
    }
    F b = (bi != UNUSED) ? val[bi] : F(0);
    reverse_partials<F>(op, a, b, r, da, db);
    E[i] = m_abs(da) * E[ai] + m_abs(r);
    T[i] = m_abs(da) * T[ai];
    if (bi != UNUSED) {
      E[i] = E[i] + m_abs(db) * E[bi];
      T[i] = T[i] + m_abs(db) * T[bi];
    }
  }
  for (uint32_t j = 0; j < t.num_outputs; ++j) val_mag[j] = E[t.output_indices[j]];
  const uint32_t seed = t.output_indices[0];
  if (c1_mag) *c1_mag = T[seed];
  std::vector<D> A(N, D{F(0), F(0)});
  A[seed] = D{F(1), F(0)};
  auto acc = [&](uint32_t k, D p, D g) {
    const F pr = m_abs(p.re), pe = m_abs(p.eps);
    A[k].re = A[k].re + pr * g.re;
    A[k].eps = A[k].eps + pr * g.eps + pe * g.re;
  };
  for (int64_t ii = (int64_t)N - 1; ii >= 0; --ii) {
    uint32_t i = (uint32_t)ii;
    uint32_t op = t.opcodes[i];
    if (op == OP_INPUT || op == OP_CONST) continue;
    D g = A[i];
    if (g.re == F(0) && g.eps == F(0)) continue;
    uint32_t ai = t.arg0[i], bi = t.arg1[i];
    D a{val[ai], T[ai]}, r{val[i], T[i]};
    if (op == OP_POWI) {
      acc(ai, powi_reverse_partial<D>((int32_t)bi, a, r), g);
      continue;
    }
    D b = (bi != UNUSED) ? D{val[bi], T[bi]} : D{F(0), F(0)};
    D da, db;
    reverse_partials<D>(op, a, b, r, da, db);
    acc(ai, da, g);
    if (bi != UNUSED) acc(bi, db, g);
  }
  F c2 = F(0);
  for (uint32_t i = 0; i < t.num_inputs; ++i) {
    if (grad_mag) grad_mag[i] = A[i].re;
    if (hv_mag) hv_mag[i] = A[i].eps;
    c2 = c2 + m_abs(v[i]) * A[i].eps;
  }
  if (c2_mag) *c2_mag = c2 / F(2);
}

// taylor_jet_2nd (src/stde/jet.rs:26-47) generalised to every output.
template <class F, int K> inline void taylor_jets(const TapeView& t, const F* x, const F* v, F* out /*[m][K]*/) {
  std::vector<Taylor<F, K>> in(t.num_inputs), buf;
  for (uint32_t i = 0; i < t.num_inputs; ++i) {
    in[i].c[0] = x[i];
    if (K > 1) in[i].c[1] = v[i];
    for (int k = 2; k < K; ++k) in[i].c[k] = F(0);
  }
  forward_tangent<Taylor<F, K>>(t, in.data(), buf);
  for (uint32_t j = 0; j < t.num_outputs; ++j) {
    const Taylor<F, K>& o = buf[t.output_indices[j]];
    for (int k = 0; k < K; ++k) out[j * K + k] = o.c[k];
  }
}

// WelfordAccumulator (src/stde/types.rs:44-113)
template <class F> struct Welford {
  F mean = F(0), m2 = F(0);
  uint64_t count = 0;
  void update(F s) {
    if (!m_isfinite(s)) return;
    count += 1;
    F k1 = (F)count;
    F delta = s - mean;
    mean = mean + delta / k1;
    F delta2 = s - mean;
    m2 = m2 + delta * delta2;
  }
  void finalize(F* est, F* var, F* se) const {
    if (count == 0) { *est = m_nan<F>(); *var = F(0); *se = F(0); return; }
    F nf = (F)count;
    if (count > 1) {
      F vv = m2 / (nf - F(1));
      if (!(vv > F(0))) vv = (vv != vv) ? F(0) : (vv < F(0) ? F(0) : vv);  // max(v, 0)
      *est = mean; *var = vv; *se = m_sqrt(vv / nf);
    } else {
      *est = mean; *var = F(0); *se = F(0);
    }
  }
};

// ---------------------------------------------------------------------------
// Structure: sparsity patterns and colorings
// (src/bytecode_tape/sparse.rs:463-776, src/sparse.rs:57-413).
// ---------------------------------------------------------------------------
inline bool sp_binary(uint32_t op) {  // is_binary_op, sparse.rs:702-718
  return op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV || op == OP_REM || op == OP_POWF ||
         op == OP_ATAN2 || op == OP_HYPOT || op == OP_MAX || op == OP_MIN;
}

// 0 linear, 1 unary nonlinear, 2 binary nonlinear, 3 zero derivative
inline int sp_class(uint32_t op) {  // classify_op, sparse.rs:643-700
  switch (op) {
    case OP_ADD: case OP_SUB: case OP_NEG: case OP_FRACT: return 0;
    case OP_MUL: case OP_DIV: case OP_POWF: case OP_ATAN2: case OP_HYPOT: return 2;
    case OP_ABS: case OP_SIGNUM: case OP_FLOOR: case OP_CEIL: case OP_ROUND: case OP_TRUNC:
    case OP_MAX: case OP_MIN: case OP_REM: return 3;
    default: return 1;
  }
}

struct DepSets {
  uint32_t n = 0, words = 0;
  std::vector<std::vector<uint64_t>> d;
  std::vector<uint32_t> bits(uint32_t i) const {  // extract_bits
    std::vector<uint32_t> r;
    for (uint32_t w = 0; w < words; ++w)
      for (uint32_t b = 0; b < 64; ++b)
        if ((d[i][w] >> b) & 1u) {
          uint32_t pos = w * 64 + b;
          if (pos < n) r.push_back(pos);
        }
    return r;
  }
  void unite(uint32_t dst, uint32_t src) {  // union_into
    if (dst == src) return;
    for (uint32_t w = 0; w < words; ++w) d[dst][w] |= d[src][w];
  }
};

// detect_sparsity_impl / detect_jacobian_sparsity_impl; `pairs` null for the
// Jacobian walk (every op propagates).
inline DepSets sp_walk(const TapeView& t, std::vector<std::pair<uint32_t, uint32_t>>* pairs) {
  DepSets D;
  D.n = t.num_inputs;
  D.words = (t.num_inputs + 63) / 64;
  D.d.assign(t.num_ops, std::vector<uint64_t>(D.words, 0));
  auto push_upper = [&](const std::vector<uint32_t>& b) {
    for (size_t ii = 0; ii < b.size(); ++ii)
      for (size_t jj = 0; jj <= ii; ++jj)
        pairs->push_back(b[ii] >= b[jj] ? std::make_pair(b[ii], b[jj]) : std::make_pair(b[jj], b[ii]));
  };
  uint32_t input_idx = 0;
  for (uint32_t i = 0; i < t.num_ops; ++i) {
    uint32_t op = t.opcodes[i];
    if (op == OP_INPUT) {
      D.d[i][input_idx / 64] |= 1ull << (input_idx % 64);
      ++input_idx;
      continue;
    }
    if (op == OP_CONST) continue;
    uint32_t a = t.arg0[i], b = t.arg1[i];
    bool bin = sp_binary(op) && b != UNUSED;
    if (!pairs) {
      D.unite(i, a);
      if (bin) D.unite(i, b);
      continue;
    }
    int cls = sp_class(op);
    if (cls == 0 || cls == 3) {
      D.unite(i, a);
      if (bin) D.unite(i, b);
    } else if (cls == 1) {
      D.unite(i, a);
      push_upper(D.bits(i));
    } else {
      std::vector<uint32_t> ba = D.bits(a), bb = D.bits(b);
      D.unite(i, a);
      D.unite(i, b);
      for (uint32_t va : ba)
        for (uint32_t vb : bb) pairs->push_back(va >= vb ? std::make_pair(va, vb) : std::make_pair(vb, va));
      if (op != OP_MUL) {
        push_upper(ba);
        push_upper(bb);
      }
    }
  }
  return D;
}

inline void hessian_pattern(const TapeView& t, std::vector<uint32_t>& rows, std::vector<uint32_t>& cols) {
  std::vector<std::pair<uint32_t, uint32_t>> p;
  sp_walk(t, &p);
  std::sort(p.begin(), p.end());
  p.erase(std::unique(p.begin(), p.end()), p.end());
  rows.clear();
  cols.clear();
  for (auto& e : p) { rows.push_back(e.first); cols.push_back(e.second); }
}

inline void jacobian_pattern(const TapeView& t, std::vector<uint32_t>& rows, std::vector<uint32_t>& cols) {
  DepSets D = sp_walk(t, nullptr);
  rows.clear();
  cols.clear();
  for (uint32_t o = 0; o < t.num_outputs; ++o)
    for (uint32_t c : D.bits(t.output_indices[o])) { rows.push_back(o); cols.push_back(c); }
}

// greedy_distance1_coloring (sparse.rs:357-413), with its u64 bitset fast path
// and the sorted-list fallback for colors >= 64.
inline uint32_t distance1_coloring(const std::vector<std::vector<uint32_t>>& adj, std::vector<uint32_t>& colors) {
  const size_t n = adj.size();
  std::vector<size_t> order(n);
  for (size_t v = 0; v < n; ++v) order[v] = v;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return adj[b].size() < adj[a].size(); });
  colors.assign(n, 0xFFFFFFFFu);
  uint32_t num = 0;
  for (size_t v : order) {
    uint64_t used = 0;
    bool fallback = false;
    for (uint32_t u : adj[v]) {
      uint32_t c = colors[u];
      if (c == 0xFFFFFFFFu) continue;
      if (c < 64) used |= 1ull << c;
      else fallback = true;
    }
    uint32_t color;
    if (!fallback) {
      uint64_t freeb = ~used;
      color = freeb ? (uint32_t)__builtin_ctzll(freeb) : 64u;
    } else {
      std::vector<uint32_t> uv;
      for (uint32_t u : adj[v])
        if (colors[u] != 0xFFFFFFFFu) uv.push_back(colors[u]);
      std::sort(uv.begin(), uv.end());
      uv.erase(std::unique(uv.begin(), uv.end()), uv.end());
      color = 0;
      for (uint32_t u : uv) {
        if (u != color) break;
        ++color;
      }
    }
    colors[v] = color;
    if (color + 1 > num) num = color + 1;
  }
  return num;
}

inline uint32_t greedy_coloring(uint32_t n, const std::vector<uint32_t>& rows, const std::vector<uint32_t>& cols,
                                std::vector<uint32_t>& colors) {
  colors.clear();
  if (n == 0) return 0;
  std::vector<std::vector<uint32_t>> adj(n), adj2(n);
  for (size_t k = 0; k < rows.size(); ++k)
    if (rows[k] != cols[k]) { adj[rows[k]].push_back(cols[k]); adj[cols[k]].push_back(rows[k]); }
  for (uint32_t v = 0; v < n; ++v) {
    for (uint32_t u : adj[v]) adj2[v].push_back(u);
    for (uint32_t u : adj[v])
      for (uint32_t w : adj[u])
        if (w != v) adj2[v].push_back(w);
    std::sort(adj2[v].begin(), adj2[v].end());
    adj2[v].erase(std::unique(adj2[v].begin(), adj2[v].end()), adj2[v].end());
  }
  return distance1_coloring(adj2, colors);
}

inline uint32_t intersection_coloring(const std::vector<uint32_t>& gk, const std::vector<uint32_t>& ck,
                                      uint32_t gdim, uint32_t cdim, std::vector<uint32_t>& colors) {
  colors.clear();
  if (cdim == 0) return 0;
  std::vector<std::vector<uint32_t>> groups(gdim), adj(cdim);
  for (size_t k = 0; k < gk.size(); ++k) groups[gk[k]].push_back(ck[k]);
  for (auto& m : groups)
    for (size_t i = 0; i < m.size(); ++i)
      for (size_t j = i + 1; j < m.size(); ++j) { adj[m[i]].push_back(m[j]); adj[m[j]].push_back(m[i]); }
  for (auto& l : adj) {
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
  }
  return distance1_coloring(adj, colors);
}

// ---------------------------------------------------------------------------
// forward_nonsmooth (src/bytecode_tape/forward.rs:87-187): one record per
// nonsmooth entry of the whole tape, in tape order.
// ---------------------------------------------------------------------------
inline bool is_nonsmooth(uint32_t op) {
  return op == OP_ABS || op == OP_MIN || op == OP_MAX || op == OP_SIGNUM || op == OP_FLOOR || op == OP_CEIL ||
         op == OP_ROUND || op == OP_TRUNC || op == OP_FRACT;
}

template <class F> inline void forward_nonsmooth(const TapeView& t, const F* x, std::vector<F>& v, std::vector<F>& sw,
                                                 std::vector<int8_t>& br) {
  forward_into<F>(t, x, v);
  sw.clear();
  br.clear();
  const F half = F(0.5);
  for (uint32_t i = 0; i < t.num_ops; ++i) {
    uint32_t op = t.opcodes[i];
    if (!is_nonsmooth(op)) continue;
    F a = v[t.arg0[i]];
    switch (op) {
      case OP_ABS:
      case OP_SIGNUM: sw.push_back(a); br.push_back(a >= F(0) ? 1 : -1); break;
      case OP_MAX: {
        F b = v[t.arg1[i]];
        sw.push_back(a - b);
        br.push_back((a >= b || b != b) ? 1 : -1);
        break;
      }
      case OP_MIN: {
        F b = v[t.arg1[i]];
        sw.push_back(a - b);
        br.push_back((a <= b || b != b) ? 1 : -1);
        break;
      }
      case OP_FLOOR:
      case OP_CEIL:
      case OP_TRUNC:
        sw.push_back(a - std::round(a));
        br.push_back(a - std::floor(a) < half ? 1 : -1);
        break;
      case OP_FRACT:
        sw.push_back(a - std::round(a));
        br.push_back(a - std::trunc(a) >= F(0) ? 1 : -1);
        break;
      case OP_ROUND: {
        F sh = a + half;
        sw.push_back(sh - std::round(sh));
        br.push_back(a - std::floor(a) < half ? 1 : -1);
        break;
      }
    }
  }
}

}  // namespace orc
