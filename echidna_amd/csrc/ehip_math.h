// Device arithmetic for the tape interpreter: scalar F, Dual<F> (forward-over-
// reverse) and Jet<F,K> (truncated Taylor series), with the reference's
// numerical conventions:
//   single-op values        src/opcode.rs:228-318
//   reverse partials        src/opcode.rs:331-511, src/kernels/mod.rs:85-285
//   Dual elementals         src/dual.rs:44-460, src/traits/std_ops.rs:14-83
//   Taylor recurrences      src/taylor_ops.rs:18-916, src/taylor.rs:49-370
// Every expression keeps the reference's evaluation order; the library is
// built with -ffp-contract=off so no a*b+c is fused.  Transcendentals come
// from the ROCm device math library (ocml), except the scalar tanh and its
// derivative (s_tanh, s_dtanh); powi uses repeated squaring like
// llvm.powi on the host.
#pragma once
// Also compiled at run time by hipRTC (tape-specialised kernels), which
// supplies the HIP runtime declarations and fixed-width integers itself.
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#else
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#define INT32_MIN (-2147483647 - 1)
#endif

namespace ehip {

enum : uint32_t {
  OP_INPUT = 0, OP_CONST = 1, OP_ADD = 2, OP_SUB = 3, OP_MUL = 4, OP_DIV = 5,
  OP_REM = 6, OP_POWF = 7, OP_ATAN2 = 8, OP_HYPOT = 9, OP_MAX = 10, OP_MIN = 11,
  OP_NEG = 12, OP_RECIP = 13, OP_SQRT = 14, OP_CBRT = 15, OP_POWI = 16,
  OP_EXP = 17, OP_EXP2 = 18, OP_EXPM1 = 19, OP_LN = 20, OP_LOG2 = 21,
  OP_LOG10 = 22, OP_LN1P = 23, OP_SIN = 24, OP_COS = 25, OP_TAN = 26,
  OP_ASIN = 27, OP_ACOS = 28, OP_ATAN = 29, OP_SINH = 30, OP_COSH = 31,
  OP_TANH = 32, OP_ASINH = 33, OP_ACOSH = 34, OP_ATANH = 35, OP_ABS = 36,
  OP_SIGNUM = 37, OP_FLOOR = 38, OP_CEIL = 39, OP_ROUND = 40, OP_TRUNC = 41,
  OP_FRACT = 42
};

#define EH_DEV static __device__ __forceinline__

// ---------------------------------------------------------------------------
// Scalar helpers (float and double overloads resolve to the *f / plain ocml
// entry points).
// ---------------------------------------------------------------------------
template <class F> EH_DEV F s_nan() { return F(__builtin_nan("")); }
template <class F> EH_DEV F s_inf() { return F(__builtin_inf()); }
template <class F> EH_DEV bool s_isnan(F x) { return x != x; }
template <class F> EH_DEV bool s_isfinite(F x) { return isfinite(x); }
template <class F> EH_DEV F s_signum(F x) { return s_isnan(x) ? x : copysign(F(1), x); }
template <class F> EH_DEV F s_fract(F x) { return x - trunc(x); }
template <class F> EH_DEV F s_powi(F a, int32_t b) {
  const bool recip = b < 0;
  F r = F(1);
  int64_t e = b;
  for (;;) {
    if (e & 1) r = r * a;
    e /= 2;
    if (e == 0) break;
    a = a * a;
  }
  return recip ? F(1) / r : r;
}
template <class F> EH_DEV bool s_exact_i32(F x, int32_t* out) {
  if (!(x > F(-2147483649.0) && x < F(2147483648.0))) return false;
  int32_t ni = (int32_t)x;
  if ((F)ni != x) return false;
  *out = ni;
  return true;
}
// tanh and its derivative, the reference's 1/cosh^2 (src/opcode.rs:463-466).
// f32: from e = exp(-2|a|) and m = expm1(-2|a|) = e - 1, tanh|a| = -m / (2 + m)
// and 1/cosh^2 a = 4e / (1 + e)^2 (the divisions by the 1-ulp hardware
// reciprocal of a value in [1, 2]): at most 3 and 4 ulp from the correctly
// rounded value over 2^20 arguments in [-50, 50] (the reference's budget
// for tanh is 16 ulp), with no cancellation near
// 0 (m, not e - 1) and no overflow (e <= 1).  About a third of the
// instructions of the device library's tanhf and coshf, which work in
// extended precision; the n=32 H=64 MLP fwd+grad went from 2.24 to 3.95 G
// points/s.  Signed zeros, infinities and NaN follow tanh and 1/cosh^2; where
// cosh^2 overflows (|a| > 44.4) the derivative is the subnormal value
// instead of 0.
// e = exp(y) and m = expm1(y) for y <= 0 (or NaN) from one range reduction:
// y = k ln2 + r with |r| <= ln2 / 2 (Cody-Waite, two-part ln2), p = expm1(r)
// by its Taylor series to r^7 (truncation below 2e-8 relative), then
// e = 2^k (1 + p) and m = 2^k p + (2^k - 1), which has no cancellation: for
// k = 0 it is p, for k < 0 the two terms add with the same sign.  y below
// -200 gives e = 0, m = -1; NaN propagates.
EH_DEV void exp_pair_f32(float y, float& e, float& m) {
  y = (y < -200.0f) ? -200.0f : y;
  const float kf = rintf(y * 1.44269504f);
  float r = fmaf(-kf, 0.693145752f, y);
  r = fmaf(-kf, 1.42860677e-6f, r);
  float p = fmaf(r, 1.98412698e-4f, 1.38888889e-3f);
  p = fmaf(r, p, 8.33333377e-3f);
  p = fmaf(r, p, 4.16666679e-2f);
  p = fmaf(r, p, 1.66666672e-1f);
  p = fmaf(r, p, 0.5f);
  p = r * p;
  p = fmaf(r, p, r);  // r + r^2 (1/2 + r/6 + ..)
  const int k = (int)kf;
  e = ldexpf(1.0f + p, k);
  m = ldexpf(p, k) + (ldexpf(1.0f, k) - 1.0f);
}
// f64: the same pair in binary64, e = exp(-2|a|) and
// m = expm1(-2|a|) from y = k ln2 + r (Cody-Waite, the two-part ln2 of
// fdlibm), p = expm1(r) by its Taylor series to r^13 (truncation below
// 1e-17 relative), e = 2^k (1 + p), m = 2^k p + (2^k - 1).  Both results
// are quotients by 2 + m = 1 + e in [1, 2]: the hardware estimate r of 1 / d
// (d = 2 + m rounded) is refined by two Newton steps, and each quotient
// q0 = n r is corrected by its residual n - (2 + m) q0, in which
// 2 + m = d + ((2 - d) + m) exactly:
//   tanh|a| = -m / (2 + m), 1/cosh^2 a = (2 / (2 + m)) (2 e / (2 + m)).
// Over |a| <= 20 they are within 1.6 and 2.8 ulp of the correctly rounded
// values (binary64 emulation of these steps against 200-bit mpmath, 4,000
// arguments; the device test bounds them at 2 and 3 ulp).
EH_DEV void exp_pair_f64(double y, double& e, double& m) {
  const double kf = rint(y * 1.4426950408889634);
  double r = fma(-kf, 6.93147180369123816490e-01, y);
  r = fma(-kf, 1.90821492927058770002e-10, r);
  double q = 1.6059043836821614599e-10;  // 1/13!
  q = fma(r, q, 2.0876756987868098979e-09);
  q = fma(r, q, 2.5052108385441718775e-08);
  q = fma(r, q, 2.7557319223985890653e-07);
  q = fma(r, q, 2.7557319223985890653e-06);
  q = fma(r, q, 2.4801587301587301566e-05);
  q = fma(r, q, 1.9841269841269841253e-04);
  q = fma(r, q, 1.3888888888888889419e-03);
  q = fma(r, q, 8.3333333333333332177e-03);
  q = fma(r, q, 4.1666666666666664354e-02);
  q = fma(r, q, 1.6666666666666665741e-01);
  q = fma(r, q, 0.5);
  const double p = fma(r * r, q, r);  // r + r^2 (1/2 + r/6 + ..)
  const int k = (int)kf;
  e = ldexp(1.0 + p, k);
  m = ldexp(p, k) + (ldexp(1.0, k) - 1.0);
}
// 1 / d for d in [1, 2]
EH_DEV double recip_12(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  return fma(r, fma(-d, r, 1.0), r);
}
// tanh a and 1/cosh^2 a in binary64 from one reduction, without branches:
// |a| is clamped at 360 (tanh rounds to +-1 from 19.1 on), 1/cosh^2 is 0
// beyond 355.29, where the reference's cosh^2 overflows, and NaN propagates
// through both.
EH_DEV void tanh_pair_f64(double a, double& t, double& dt) {
  const double ax = fabs(a);
  double e, m;
  exp_pair_f64(-2.0 * (ax > 360.0 ? 360.0 : ax), e, m);
  const double d = 2.0 + m, derr = (2.0 - d) + m;  // 2 + m = d + derr exactly
  double r = recip_12(d);
  const double t0 = -m * r;
  t = copysign(fma(r, fma(d, -t0, -m) - derr * t0, t0), a);  // t0 + (-m - (2 + m) t0) / d
  // 1/cosh^2 = (1 + tanh|a|)(1 - tanh|a|) = (2 / (2 + m)) (2 e / (2 + m)),
  // each quotient corrected by its residual as above
  const double A0 = 2.0 * r, B0 = (2.0 * e) * r;
  const double A = fma(r, fma(d, -A0, 2.0) - derr * A0, A0);
  const double B = fma(r, fma(d, -B0, 2.0 * e) - derr * B0, B0);
  dt = ax > 355.29 ? 0.0 : A * B;
}
template <class F> EH_DEV F s_tanh(F a) {
  if constexpr (sizeof(F) == 4) {
    // 2 + m lies in [1, 2]: the hardware reciprocal (1 ulp) needs no scaling
    F e, m;
    exp_pair_f32(F(-2) * fabs(a), e, m);
    return copysign(-m * __builtin_amdgcn_rcpf(F(2) + m), a);
  } else {
    double t, dt;
    tanh_pair_f64(a, t, dt);
    return t;
  }
}
template <class F> EH_DEV F s_dtanh(F a) {
  if constexpr (sizeof(F) == 4) {
    F e, m;
    exp_pair_f32(F(-2) * fabs(a), e, m);
    const F r = __builtin_amdgcn_rcpf(F(1) + e);  // 1 + e in [1, 2]
    return (F(4) * e) * (r * r);
  } else {
    double t, dt;
    tanh_pair_f64(a, t, dt);
    return dt;
  }
}
// Both at once (one reduction in binary64), for straight-line code with
// many tanh: the binary64 pair is a called function there, so its ~30
// dependent steps are not interleaved across units (which held hundreds of
// registers and spilled), and the code is not repeated per unit.
struct TanhPair {
  double t, dt;
};
static __device__ __attribute__((noinline)) TanhPair tanh_pair_f64_call(double a) {
  TanhPair o;
  tanh_pair_f64(a, o.t, o.dt);
  return o;
}
template <class F> EH_DEV void s_tanh_dtanh(F a, F& t, F& dt) {
  if constexpr (sizeof(F) == 4) {
    t = s_tanh(a);
    dt = s_dtanh(a);
  } else {
    const TanhPair o = tanh_pair_f64_call(a);
    t = o.t;
    dt = o.dt;
  }
}
template <class F> EH_DEV F s_ln2() { return F(0.69314718055994530942); }
template <class F> EH_DEV F s_ln10() { return F(2.30258509299404568402); }

// Single-op value on plain scalars (src/opcode.rs:228-318).
template <class F> EH_DEV F s_eval(uint32_t op, F a, F b) {
  switch (op) {
    case OP_ADD: return a + b;
    case OP_SUB: return a - b;
    case OP_MUL: return a * b;
    case OP_DIV: return a / b;
    case OP_REM: return fmod(a, b);
    case OP_POWF: return pow(a, b);
    case OP_ATAN2: return atan2(a, b);
    case OP_HYPOT: return hypot(a, b);
    case OP_MAX: return (a >= b || s_isnan(b)) ? a : b;
    case OP_MIN: return (a <= b || s_isnan(b)) ? a : b;
    case OP_NEG: return -a;
    case OP_RECIP: return F(1) / a;
    case OP_SQRT: return sqrt(a);
    case OP_CBRT: return cbrt(a);
    case OP_EXP: return exp(a);
    case OP_EXP2: return exp2(a);
    case OP_EXPM1: return expm1(a);
    case OP_LN: return log(a);
    case OP_LOG2: return log2(a);
    case OP_LOG10: return log10(a);
    case OP_LN1P: return log1p(a);
    case OP_SIN: return sin(a);
    case OP_COS: return cos(a);
    case OP_TAN: return tan(a);
    case OP_ASIN: return asin(a);
    case OP_ACOS: return acos(a);
    case OP_ATAN: return atan(a);
    case OP_SINH: return sinh(a);
    case OP_COSH: return cosh(a);
    case OP_TANH: return s_tanh(a);
    case OP_ASINH: return asinh(a);
    case OP_ACOSH: return acosh(a);
    case OP_ATANH: return atanh(a);
    case OP_ABS: return fabs(a);
    case OP_SIGNUM: return s_signum(a);
    case OP_FLOOR: return floor(a);
    case OP_CEIL: return ceil(a);
    case OP_ROUND: return round(a);
    case OP_TRUNC: return trunc(a);
    case OP_FRACT: return s_fract(a);
    default: return s_nan<F>();
  }
}

// Derivative helpers (src/kernels/mod.rs:85-285) on plain scalars.
template <class F> EH_DEV F k_ln_deriv(F a) { return (a >= F(0)) ? F(1) / a : s_nan<F>(); }
template <class F> EH_DEV F k_log2_deriv(F a) { return (a >= F(0)) ? F(1) / (a * s_ln2<F>()) : s_nan<F>(); }
template <class F> EH_DEV F k_log10_deriv(F a) { return (a >= F(0)) ? F(1) / (a * s_ln10<F>()) : s_nan<F>(); }
template <class F> EH_DEV F k_ln1p_deriv(F a) { return (a >= -F(1)) ? F(1) / (F(1) + a) : s_nan<F>(); }
template <class F> EH_DEV F k_atanh_deriv(F a) {
  return (a >= -F(1) && a <= F(1)) ? F(1) / ((F(1) - a) * (F(1) + a)) : s_nan<F>();
}
template <class F> EH_DEV F k_abs_deriv(F a) { return (a == F(0)) ? F(0) : s_signum(a); }
template <class F> EH_DEV F k_atan_deriv(F a) {
  if (fabs(a) > F(1e8)) {
    F inv = F(1) / a;
    return inv * inv / (F(1) + inv * inv);
  }
  return F(1) / (F(1) + a * a);
}
template <class F> EH_DEV F k_asinh_deriv(F a) {
  if (fabs(a) > F(1e8)) {
    F inv = F(1) / a;
    return fabs(inv) / sqrt(F(1) + inv * inv);
  }
  return F(1) / sqrt(a * a + F(1));
}
template <class F> EH_DEV F k_acosh_deriv(F a) {
  if (a < F(1)) return s_nan<F>();
  if (a > F(1e8)) {
    F inv = F(1) / a;
    return fabs(inv) / sqrt(F(1) - inv * inv);
  }
  return F(1) / sqrt((a - F(1)) * (a + F(1)));
}

// Reverse partials on plain scalars (src/opcode.rs:331-483).  `ea`/`eb`
// say which partial is consumed; the other is left at zero.
template <class F> EH_DEV void s_partials(uint32_t op, F a, F b, F r, F& da, F& db) {
  da = F(0);
  db = F(0);
  switch (op) {
    case OP_ADD: da = F(1); db = F(1); return;
    case OP_SUB: da = F(1); db = F(-1); return;
    case OP_MUL: da = b; db = a; return;
    case OP_DIV: { F inv = F(1) / b; da = inv; db = -r * inv; return; }
    case OP_REM: da = F(1); db = -trunc(a / b); return;
    case OP_POWF: {
      if (b == F(0)) { db = (a > F(0)) ? log(a) : F(0); return; }
      if (a == F(0) || r == F(0) || !s_isfinite(a) || !s_isfinite(r)) da = b * pow(a, b - F(1));
      else da = b * r / a;
      db = (r == F(0) || a <= F(0)) ? F(0) : r * log(a);
      return;
    }
    case OP_ATAN2: {
      F h = hypot(a, b);
      if (h == F(0)) return;
      da = b / h / h;
      db = -a / h / h;
      return;
    }
    case OP_HYPOT:
      if (r == F(0)) return;
      da = a / r;
      db = b / r;
      return;
    case OP_MAX: if (a >= b || s_isnan(b)) da = F(1); else db = F(1); return;
    case OP_MIN: if (a <= b || s_isnan(b)) da = F(1); else db = F(1); return;
    case OP_NEG: da = F(-1); return;
    case OP_RECIP: { F inv = F(1) / a; da = -inv * inv; return; }
    case OP_SQRT: { F two = F(1) + F(1); da = F(1) / (two * r); return; }
    case OP_CBRT: da = F(1) / (F(3) * r * r); return;
    case OP_EXP: da = r; return;
    case OP_EXP2: da = r * s_ln2<F>(); return;
    case OP_EXPM1: da = r + F(1); return;
    case OP_LN: da = k_ln_deriv(a); return;
    case OP_LOG2: da = k_log2_deriv(a); return;
    case OP_LOG10: da = k_log10_deriv(a); return;
    case OP_LN1P: da = k_ln1p_deriv(a); return;
    case OP_SIN: da = cos(a); return;
    case OP_COS: da = -sin(a); return;
    case OP_TAN: { F c = cos(a); da = F(1) / (c * c); return; }
    case OP_ASIN: da = F(1) / sqrt((F(1) - a) * (F(1) + a)); return;
    case OP_ACOS: da = -F(1) / sqrt((F(1) - a) * (F(1) + a)); return;
    case OP_ATAN: da = k_atan_deriv(a); return;
    case OP_SINH: da = cosh(a); return;
    case OP_COSH: da = sinh(a); return;
    case OP_TANH: da = s_dtanh(a); return;
    case OP_ASINH: da = k_asinh_deriv(a); return;
    case OP_ACOSH: da = k_acosh_deriv(a); return;
    case OP_ATANH: da = k_atanh_deriv(a); return;
    case OP_ABS: da = k_abs_deriv(a); return;
    case OP_FRACT: da = F(1); return;
    default: return;  // signum/floor/ceil/round/trunc: (0, 0)
  }
}

// powi partial (src/opcode.rs:503-511)
template <class F> EH_DEV F s_powi_partial(int32_t e, F a, F r) {
  if (e == 0) return F(0);
  if (e == INT32_MIN) return F(e) * r / a;
  return F(e) * s_powi(a, e - 1);
}

// ---------------------------------------------------------------------------
// Dual<F>
// ---------------------------------------------------------------------------
template <class F> struct Dual { F re, eps; };
template <class F> EH_DEV Dual<F> dc(F v) { return Dual<F>{v, F(0)}; }
template <class F> EH_DEV Dual<F> d_add(Dual<F> a, Dual<F> b) { return {a.re + b.re, a.eps + b.eps}; }
template <class F> EH_DEV Dual<F> d_sub(Dual<F> a, Dual<F> b) { return {a.re - b.re, a.eps - b.eps}; }
template <class F> EH_DEV Dual<F> d_mul(Dual<F> a, Dual<F> b) { return {a.re * b.re, a.re * b.eps + a.eps * b.re}; }
template <class F> EH_DEV Dual<F> d_div(Dual<F> a, Dual<F> b) {
  F inv = F(1) / b.re;
  F re = a.re / b.re;
  return {re, (a.eps - re * b.eps) * inv};
}
template <class F> EH_DEV Dual<F> d_neg(Dual<F> a) { return {-a.re, -a.eps}; }
template <class F> EH_DEV bool d_zero(Dual<F> a) { return a.re == F(0) && a.eps == F(0); }
template <class F> EH_DEV Dual<F> d_chain(Dual<F> x, F f, F df) {
  return {f, (x.eps == F(0)) ? F(0) : x.eps * df};
}
// Dual tanh in f32 may use s_tanh / s_dtanh when no lane of the wave has
// |x| > 40 or a non-finite x (a wave-uniform test).  There cosh^2 cannot
// overflow, every value the reference's dual formulas produce is finite, and
// the cheaper forms give the same NaN / Inf pattern for every tangent; past
// it the wave runs the reference's formulas.
template <class F> EH_DEV bool d_tanh_fast(F x) {
  if constexpr (sizeof(F) == 4) return !__any(!(fabs(x) <= F(40)));
  else return false;
}
template <class F> EH_DEV Dual<F> d_powi(Dual<F> x, int32_t n) {
  if (n == 0) return dc(F(1));
  F val = s_powi(x.re, n);
  F deriv = (n == INT32_MIN) ? F(n) * val / x.re : F(n) * s_powi(x.re, n - 1);
  return d_chain(x, val, deriv);
}
template <class F> EH_DEV Dual<F> d_powf(Dual<F> x, Dual<F> n) {
  if (n.eps == F(0)) {
    int32_t ni;
    if (s_exact_i32(n.re, &ni)) return d_powi(x, ni);
  }
  if (n.re == F(0)) {
    F dy = (x.re > F(0)) ? log(x.re) : F(0);
    return {F(1), dy * n.eps};
  }
  F val = pow(x.re, n.re);
  F dx;
  if (x.eps == F(0)) dx = F(0);
  else if (x.re == F(0) || val == F(0) || !s_isfinite(x.re) || !s_isfinite(val))
    dx = n.re * pow(x.re, n.re - F(1)) * x.eps;
  else
    dx = n.re * val / x.re * x.eps;
  F dy = (n.eps == F(0) || val == F(0) || x.re <= F(0)) ? F(0) : val * log(x.re) * n.eps;
  return {val, dx + dy};
}
template <class F> EH_DEV Dual<F> d_hypot(Dual<F> x, Dual<F> y) {
  F h = hypot(x.re, y.re);
  if (h == F(0)) return dc(h);
  F da = x.re / h, db = y.re / h;
  return {h, da * x.eps + db * y.eps};
}
template <class F> EH_DEV Dual<F> d_atan2(Dual<F> x, Dual<F> y) {
  F h = hypot(x.re, y.re);
  F ds = F(0), dt = F(0);
  if (h != F(0)) { ds = y.re / h / h; dt = -x.re / h / h; }
  return {atan2(x.re, y.re), ds * x.eps + dt * y.eps};
}

// Dual value of one op (eval_forward at T = Dual<F>).
template <class F> EH_DEV Dual<F> d_eval(uint32_t op, Dual<F> a, Dual<F> b) {
  F x = a.re;
  switch (op) {
    case OP_ADD: return d_add(a, b);
    case OP_SUB: return d_sub(a, b);
    case OP_MUL: return d_mul(a, b);
    case OP_DIV: return d_div(a, b);
    case OP_REM: return {fmod(a.re, b.re), a.eps - b.eps * trunc(a.re / b.re)};
    case OP_POWF: return d_powf(a, b);
    case OP_ATAN2: return d_atan2(a, b);
    case OP_HYPOT: return d_hypot(a, b);
    case OP_MAX: return (a.re >= b.re || s_isnan(b.re)) ? a : b;
    case OP_MIN: return (a.re <= b.re || s_isnan(b.re)) ? a : b;
    case OP_NEG: return d_neg(a);
    case OP_RECIP: { F inv = F(1) / x; return d_chain(a, inv, -inv * inv); }
    case OP_SQRT: { F s = sqrt(x); F two = F(1) + F(1); return d_chain(a, s, F(1) / (two * s)); }
    case OP_CBRT: { F c = cbrt(x); return d_chain(a, c, F(1) / (F(3) * c * c)); }
    case OP_EXP: { F e = exp(x); return d_chain(a, e, e); }
    case OP_EXP2: { F e = exp2(x); return d_chain(a, e, e * s_ln2<F>()); }
    case OP_EXPM1: return d_chain(a, expm1(x), exp(x));
    case OP_LN: return d_chain(a, log(x), k_ln_deriv(x));
    case OP_LOG2: return d_chain(a, log2(x), k_log2_deriv(x));
    case OP_LOG10: return d_chain(a, log10(x), k_log10_deriv(x));
    case OP_LN1P: return d_chain(a, log1p(x), k_ln1p_deriv(x));
    case OP_SIN: return d_chain(a, sin(x), cos(x));
    case OP_COS: return d_chain(a, cos(x), -sin(x));
    case OP_TAN: { F c = cos(x); return d_chain(a, tan(x), F(1) / (c * c)); }
    case OP_ASIN: return d_chain(a, asin(x), F(1) / sqrt((F(1) - x) * (F(1) + x)));
    case OP_ACOS: return d_chain(a, acos(x), -F(1) / sqrt((F(1) - x) * (F(1) + x)));
    case OP_ATAN: return d_chain(a, atan(x), k_atan_deriv(x));
    case OP_SINH: return d_chain(a, sinh(x), cosh(x));
    case OP_COSH: return d_chain(a, cosh(x), sinh(x));
    case OP_TANH: {
      if (d_tanh_fast(x)) return d_chain(a, s_tanh(x), s_dtanh(x));
      F c = cosh(x);
      return d_chain(a, tanh(x), F(1) / (c * c));
    }
    case OP_ASINH: return d_chain(a, asinh(x), k_asinh_deriv(x));
    case OP_ACOSH: return d_chain(a, acosh(x), k_acosh_deriv(x));
    case OP_ATANH: return d_chain(a, atanh(x), k_atanh_deriv(x));
    case OP_ABS: return d_chain(a, fabs(x), k_abs_deriv(x));
    case OP_SIGNUM: return dc(s_signum(x));
    case OP_FLOOR: return dc(floor(x));
    case OP_CEIL: return dc(ceil(x));
    case OP_ROUND: return dc(round(x));
    case OP_TRUNC: return dc(trunc(x));
    case OP_FRACT: return {s_fract(x), a.eps};
    default: return dc(s_nan<F>());
  }
}

// Reverse partials at T = Dual<F> (reverse_partials::<Dual<F>>).
template <class F> EH_DEV void d_partials(uint32_t op, Dual<F> a, Dual<F> b, Dual<F> r, Dual<F>& da,
                                          Dual<F>& db) {
  const Dual<F> zero = dc(F(0)), one = dc(F(1));
  da = zero;
  db = zero;
  switch (op) {
    case OP_ADD: da = one; db = one; return;
    case OP_SUB: da = one; db = d_neg(one); return;
    case OP_MUL: da = b; db = a; return;
    case OP_DIV: { Dual<F> inv = d_div(one, b); da = inv; db = d_mul(d_neg(r), inv); return; }
    case OP_REM: da = one; db = d_neg(dc(trunc(a.re / b.re))); return;
    case OP_POWF: {
      if (b.re == F(0)) { db = (a.re > F(0)) ? d_eval(OP_LN, a, a) : zero; return; }
      if (a.re == F(0) || r.re == F(0) || !s_isfinite(a.re) || !s_isfinite(r.re))
        da = d_mul(b, d_powf(a, d_sub(b, one)));
      else
        da = d_div(d_mul(b, r), a);
      db = (r.re == F(0) || a.re <= F(0)) ? zero : d_mul(r, d_eval(OP_LN, a, a));
      return;
    }
    case OP_ATAN2: {
      Dual<F> h = d_hypot(a, b);
      if (h.re == F(0)) return;
      da = d_div(d_div(b, h), h);
      db = d_div(d_div(d_neg(a), h), h);
      return;
    }
    case OP_HYPOT:
      if (r.re == F(0)) return;
      da = d_div(a, r);
      db = d_div(b, r);
      return;
    case OP_MAX: if (a.re >= b.re || s_isnan(b.re)) da = one; else db = one; return;
    case OP_MIN: if (a.re <= b.re || s_isnan(b.re)) da = one; else db = one; return;
    case OP_NEG: da = d_neg(one); return;
    case OP_RECIP: { Dual<F> inv = d_div(one, a); da = d_mul(d_neg(inv), inv); return; }
    case OP_SQRT: { Dual<F> two = d_add(one, one); da = d_div(one, d_mul(two, r)); return; }
    case OP_CBRT: da = d_div(one, d_mul(d_mul(dc(F(3)), r), r)); return;
    case OP_EXP: da = r; return;
    case OP_EXP2: da = d_mul(r, dc(s_ln2<F>())); return;
    case OP_EXPM1: da = d_add(r, one); return;
    case OP_LN: da = (a.re >= F(0)) ? d_div(one, a) : dc(s_nan<F>()); return;
    case OP_LOG2: da = (a.re >= F(0)) ? d_div(one, d_mul(a, dc(s_ln2<F>()))) : dc(s_nan<F>()); return;
    case OP_LOG10: da = (a.re >= F(0)) ? d_div(one, d_mul(a, dc(s_ln10<F>()))) : dc(s_nan<F>()); return;
    case OP_LN1P: da = (a.re >= -F(1)) ? d_div(one, d_add(one, a)) : dc(s_nan<F>()); return;
    case OP_SIN: da = d_eval(OP_COS, a, a); return;
    case OP_COS: da = d_neg(d_eval(OP_SIN, a, a)); return;
    case OP_TAN: { Dual<F> c = d_eval(OP_COS, a, a); da = d_div(one, d_mul(c, c)); return; }
    case OP_ASIN: da = d_div(one, d_eval(OP_SQRT, d_mul(d_sub(one, a), d_add(one, a)), a)); return;
    case OP_ACOS: da = d_div(d_neg(one), d_eval(OP_SQRT, d_mul(d_sub(one, a), d_add(one, a)), a)); return;
    case OP_ATAN:
      if (fabs(a.re) > F(1e8)) {
        Dual<F> inv = d_div(one, a);
        da = d_div(d_mul(inv, inv), d_add(one, d_mul(inv, inv)));
      } else {
        da = d_div(one, d_add(one, d_mul(a, a)));
      }
      return;
    case OP_SINH: da = d_eval(OP_COSH, a, a); return;
    case OP_COSH: da = d_eval(OP_SINH, a, a); return;
    case OP_TANH: {
      if (d_tanh_fast(a.re)) {
        // 1/cosh^2 and its tangent -2 tanh / cosh^2 * eps; a zero tangent
        // stays +0 as in the reference's dual division
        const F t = s_tanh(a.re), d = s_dtanh(a.re);
        da = {d, (a.eps == F(0)) ? F(0) : -(F(2) * t * d) * a.eps};
        return;
      }
      Dual<F> c = d_eval(OP_COSH, a, a);
      da = d_div(one, d_mul(c, c));
      return;
    }
    case OP_ASINH:
      if (fabs(a.re) > F(1e8)) {
        Dual<F> inv = d_div(one, a);
        da = d_div(d_eval(OP_ABS, inv, inv), d_eval(OP_SQRT, d_add(one, d_mul(inv, inv)), inv));
      } else {
        da = d_div(one, d_eval(OP_SQRT, d_add(d_mul(a, a), one), a));
      }
      return;
    case OP_ACOSH:
      if (a.re < F(1)) { da = dc(s_nan<F>()); return; }
      if (a.re > F(1e8)) {
        Dual<F> inv = d_div(one, a);
        da = d_div(d_eval(OP_ABS, inv, inv), d_eval(OP_SQRT, d_sub(one, d_mul(inv, inv)), inv));
      } else {
        da = d_div(one, d_eval(OP_SQRT, d_mul(d_sub(a, one), d_add(a, one)), a));
      }
      return;
    case OP_ATANH:
      if (a.re >= -F(1) && a.re <= F(1)) da = d_div(one, d_mul(d_sub(one, a), d_add(one, a)));
      else da = dc(s_nan<F>());
      return;
    case OP_ABS: da = (a.re == F(0)) ? zero : dc(s_signum(a.re)); return;
    case OP_FRACT: da = one; return;
    default: return;
  }
}
template <class F> EH_DEV Dual<F> d_powi_partial(int32_t e, Dual<F> a, Dual<F> r) {
  if (e == 0) return dc(F(0));
  if (e == INT32_MIN) return d_div(d_mul(dc(F(e)), r), a);
  return d_mul(dc(F(e)), d_powi(a, e - 1));
}

// ---------------------------------------------------------------------------
// Jet<F,K>: truncated Taylor series, c[k] = f^(k)/k! (src/taylor_ops.rs).
// ---------------------------------------------------------------------------
template <class F, int K> struct Jet { F c[K]; };

namespace jt {
template <class F, int K> EH_DEV void mul(const F* a, const F* b, F* c) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    F s = F(0);
#pragma unroll
    for (int j = 0; j <= k; ++j) s = s + a[j] * b[k - j];
    c[k] = s;
  }
}
template <class F, int K> EH_DEV void div(const F* a, const F* b, F* c) {
  c[0] = a[0] / b[0];
  F inv = F(1) / b[0];
#pragma unroll
  for (int k = 1; k < K; ++k) {
    F s = a[k];
#pragma unroll
    for (int j = 1; j <= k; ++j) s = s - b[j] * c[k - j];
    c[k] = s * inv;
  }
}
template <class F, int K> EH_DEV void recip(const F* a, F* c) {
  F inv = F(1) / a[0];
  c[0] = inv;
#pragma unroll
  for (int k = 1; k < K; ++k) {
    F s = F(0);
#pragma unroll
    for (int j = 1; j <= k; ++j) s = s + a[j] * c[k - j];
    c[k] = -s * inv;
  }
}
template <class F, int K> EH_DEV void scale(const F* a, F s, F* c) {
#pragma unroll
  for (int k = 0; k < K; ++k) c[k] = s * a[k];
}
// exp_ with exp(a[0]) given (a batch GEMM forms it once per shared primal)
template <class F, int K> EH_DEV void exp_tail(const F* a, F e0, F* c) {
  c[0] = e0;
#pragma unroll
  for (int k = 1; k < K; ++k) {
    F s = F(0);
#pragma unroll
    for (int j = 1; j <= k; ++j) s = s + F(j) * a[j] * c[k - j];
    c[k] = s / F(k);
  }
}
template <class F, int K> EH_DEV void exp_(const F* a, F* c) { exp_tail<F, K>(a, exp(a[0]), c); }
template <class F, int K> EH_DEV void nanj(F* c) {
#pragma unroll
  for (int k = 0; k < K; ++k) c[k] = s_nan<F>();
}
template <class F, int K> EH_DEV void vert(F* c) {
  c[0] = F(0);
#pragma unroll
  for (int k = 1; k < K; ++k) c[k] = s_inf<F>();
}
template <class F, int K> EH_DEV void ln_(const F* a, F* c) {
  if (a[0] < F(0)) { nanj<F, K>(c); return; }
  F inv = F(1) / a[0];
  c[0] = log(a[0]);
#pragma unroll
  for (int k = 1; k < K; ++k) {
    F s = F(0);
#pragma unroll
    for (int j = 1; j < k; ++j) s = s + F(j) * c[j] * a[k - j];
    c[k] = (a[k] - s / F(k)) * inv;
  }
}
template <class F, int K> EH_DEV void sqrt_(const F* a, F* c) {
  if (a[0] == F(0)) { vert<F, K>(c); return; }
  if (a[0] < F(0)) { nanj<F, K>(c); return; }
  c[0] = sqrt(a[0]);
  F two_c0 = F(2) * c[0];
#pragma unroll
  for (int k = 1; k < K; ++k) {
    F s = F(0);
#pragma unroll
    for (int j = 1; j < k; ++j) s = s + c[j] * c[k - j];
    c[k] = (a[k] - s) / two_c0;
  }
}
// sin_cos with sin(a[0]) and cos(a[0]) given
template <class F, int K> EH_DEV void sin_cos_tail(const F* a, F s0, F c0, F* s, F* co) {
  s[0] = s0;
  co[0] = c0;
#pragma unroll
  for (int k = 1; k < K; ++k) {
    F inv_k = F(1) / F(k);
    F ss = F(0), sc = F(0);
#pragma unroll
    for (int j = 1; j <= k; ++j) {
      ss = ss + F(j) * a[j] * co[k - j];
      sc = sc + F(j) * a[j] * s[k - j];
    }
    s[k] = ss * inv_k;
    co[k] = -sc * inv_k;
  }
}
template <class F, int K> EH_DEV void sin_cos(const F* a, F* s, F* co) {
  sin_cos_tail<F, K>(a, sin(a[0]), cos(a[0]), s, co);
}
template <class F, int K> EH_DEV void sinh_cosh(const F* a, F* sh, F* ch) {
  sh[0] = sinh(a[0]);
  ch[0] = cosh(a[0]);
#pragma unroll
  for (int k = 1; k < K; ++k) {
    F inv_k = F(1) / F(k);
    F ss = F(0), sc = F(0);
#pragma unroll
    for (int j = 1; j <= k; ++j) {
      ss = ss + F(j) * a[j] * ch[k - j];
      sc = sc + F(j) * a[j] * sh[k - j];
    }
    sh[k] = ss * inv_k;
    ch[k] = sc * inv_k;
  }
}
template <class F, int K> EH_DEV void integ(const F* a, const F* g, F* c) {
#pragma unroll
  for (int k = 1; k < K; ++k) {
    F s = F(0);
#pragma unroll
    for (int j = 1; j <= k; ++j) s = s + F(j) * a[j] * g[k - j];
    c[k] = s / F(k);
  }
}
template <class F, int K> EH_DEV void atan_(const F* a, F* c) {
  F s1[K], s2[K];
  mul<F, K>(a, a, s1);
#pragma unroll
  for (int k = 0; k < K; ++k) s2[k] = s1[k];
  s2[0] = F(1) + s1[0];
  c[0] = atan(a[0]);
  recip<F, K>(s2, s1);
  integ<F, K>(a, s1, c);
}
template <class F, int K> EH_DEV void asin_(const F* a, F* c) {
  F s1[K], s2[K];
  c[0] = asin(a[0]);
  mul<F, K>(a, a, s1);
  s2[0] = (F(1) - a[0]) * (F(1) + a[0]);
#pragma unroll
  for (int k = 1; k < K; ++k) s2[k] = -s1[k];
  sqrt_<F, K>(s2, s1);
  recip<F, K>(s1, s2);
  integ<F, K>(a, s2, c);
}
template <class F, int K> EH_DEV void tan_(const F* a, F* c) {
  F sc[K];
  c[0] = tan(a[0]);
  sc[0] = F(1) + c[0] * c[0];
#pragma unroll
  for (int k = 1; k < K; ++k) {
    F s = F(0);
#pragma unroll
    for (int j = 1; j <= k; ++j) s = s + F(j) * a[j] * sc[k - j];
    c[k] = s / F(k);
    F sk = F(0);
#pragma unroll
    for (int j = 0; j <= k; ++j) sk = sk + c[j] * c[k - j];
    sc[k] = sk;
  }
}
// tanh_ with tanh(a[0]) given
template <class F, int K> EH_DEV void tanh_tail(const F* a, F t0, F* c) {
  F sc[K];
  c[0] = t0;
  sc[0] = F(1) - c[0] * c[0];
#pragma unroll
  for (int k = 1; k < K; ++k) {
    F s = F(0);
#pragma unroll
    for (int j = 1; j <= k; ++j) s = s + F(j) * a[j] * sc[k - j];
    c[k] = s / F(k);
    F sk = F(0);
#pragma unroll
    for (int j = 0; j <= k; ++j) sk = sk + c[j] * c[k - j];
    sc[k] = -sk;
  }
}
template <class F, int K> EH_DEV void tanh_(const F* a, F* c) { tanh_tail<F, K>(a, tanh(a[0]), c); }
template <class F, int K> EH_DEV void asinh_(const F* a, F* c) {
  F s1[K], s2[K];
  c[0] = asinh(a[0]);
  mul<F, K>(a, a, s1);
#pragma unroll
  for (int k = 0; k < K; ++k) s2[k] = s1[k];
  s2[0] = F(1) + s1[0];
  sqrt_<F, K>(s2, s1);
  recip<F, K>(s1, s2);
  integ<F, K>(a, s2, c);
}
template <class F, int K> EH_DEV void acosh_(const F* a, F* c) {
  if (a[0] < F(1)) { nanj<F, K>(c); return; }
  F s1[K], s2[K];
  c[0] = acosh(a[0]);
  mul<F, K>(a, a, s1);
#pragma unroll
  for (int k = 0; k < K; ++k) s2[k] = s1[k];
  s2[0] = (a[0] - F(1)) * (a[0] + F(1));
  sqrt_<F, K>(s2, s1);
  recip<F, K>(s1, s2);
  integ<F, K>(a, s2, c);
}
template <class F, int K> EH_DEV void atanh_(const F* a, F* c) {
  if (a[0] < -F(1) || a[0] > F(1)) { nanj<F, K>(c); return; }
  F s1[K], s2[K];
  c[0] = atanh(a[0]);
  mul<F, K>(a, a, s1);
  s2[0] = (F(1) - a[0]) * (F(1) + a[0]);
#pragma unroll
  for (int k = 1; k < K; ++k) s2[k] = -s1[k];
  recip<F, K>(s2, s1);
  integ<F, K>(a, s1, c);
}
template <class F, int K> EH_DEV void powi_sq(const F* a, int32_t n, F* c) {
  F s1[K], s2[K];
  uint32_t power = (n < 0) ? (uint32_t)(-(int64_t)n) : (uint32_t)n;
  c[0] = F(1);
#pragma unroll
  for (int k = 1; k < K; ++k) c[k] = F(0);
#pragma unroll
  for (int k = 0; k < K; ++k) s1[k] = a[k];
  while (power > 0) {
    if (power & 1u) {
      mul<F, K>(c, s1, s2);
#pragma unroll
      for (int k = 0; k < K; ++k) c[k] = s2[k];
    }
    power >>= 1;
    if (power > 0) {
      mul<F, K>(s1, s1, s2);
#pragma unroll
      for (int k = 0; k < K; ++k) s1[k] = s2[k];
    }
  }
  if (n < 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) s1[k] = c[k];
    recip<F, K>(s1, c);
  }
}
template <class F, int K> EH_DEV void powi_(const F* a, int32_t n, F* c) {
  if (n == 0) {
    c[0] = F(1);
#pragma unroll
    for (int k = 1; k < K; ++k) c[k] = F(0);
    return;
  }
  if (n == 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) c[k] = a[k];
    return;
  }
  if (n == -1) { recip<F, K>(a, c); return; }
  uint32_t abs_n = (n < 0) ? (uint32_t)(-(int64_t)n) : (uint32_t)n;
  if (a[0] <= F(0) || abs_n <= 8u) {
    powi_sq<F, K>(a, n, c);
  } else {
    F s1[K], s2[K];
    ln_<F, K>(a, s1);
    scale<F, K>(s1, F(n), s2);
    exp_<F, K>(s2, c);
    c[0] = s_powi(a[0], n);
  }
}
template <class F, int K> EH_DEV void powf_(const F* a, const F* b, F* c) {
  bool const_b = true;
#pragma unroll
  for (int k = 1; k < K; ++k)
    if (b[k] != F(0)) const_b = false;
  int32_t ni;
  if (const_b && s_exact_i32(b[0], &ni)) { powi_<F, K>(a, ni, c); return; }
  if (a[0] == F(0)) {
    if (s_exact_i32(b[0], &ni)) { nanj<F, K>(c); return; }
    c[0] = pow(a[0], b[0]);
#pragma unroll
    for (int k = 1; k < K; ++k) c[k] = (F(k) < b[0]) ? F(0) : s_inf<F>();
    return;
  }
  F s1[K], s2[K];
  ln_<F, K>(a, s1);
  mul<F, K>(b, s1, s2);
  exp_<F, K>(s2, c);
  c[0] = (a[0] < F(0)) ? s_nan<F>() : pow(a[0], b[0]);
}
template <class F, int K> EH_DEV void cbrt_(const F* a, F* c) {
  if (a[0] == F(0)) { vert<F, K>(c); return; }
  F s1[K], s2[K], t[K];
  F third = F(1) / F(3);
  bool neg = a[0] < F(0);
#pragma unroll
  for (int i = 0; i < K; ++i) t[i] = neg ? -a[i] : a[i];
  ln_<F, K>(t, s1);
  scale<F, K>(s1, third, s2);
  exp_<F, K>(s2, c);
  c[0] = cbrt(a[0]);
  if (neg) {
#pragma unroll
    for (int i = 1; i < K; ++i) c[i] = -c[i];
  }
}
template <class F, int K> EH_DEV void hypot_(const F* a, const F* b, F* c) {
  int m = 0;
  F am0 = fabs(a[0]), bm0 = fabs(b[0]);
  F mx = s_isnan(am0) ? bm0 : (s_isnan(bm0) ? am0 : (am0 > bm0 ? am0 : bm0));
  if (mx == F(0)) {
    if (s_isnan(a[0]) || s_isnan(b[0])) { nanj<F, K>(c); return; }
    int found = -1;
#pragma unroll
    for (int k = K - 1; k >= 1; --k)
      if (a[k] != F(0) || b[k] != F(0)) found = k;
    if (found < 0) {
#pragma unroll
      for (int k = 0; k < K; ++k) c[k] = F(0);
      return;
    }
    m = found;
    F am = F(0), bm = F(0);
#pragma unroll
    for (int k = 1; k < K; ++k)
      if (k == m) { am = a[k]; bm = b[k]; }
    if (s_isnan(am) || s_isnan(bm)) {
#pragma unroll
      for (int k = 0; k < K; ++k) c[k] = (k < m) ? F(0) : ((k == m) ? hypot(am, bm) : s_nan<F>());
      return;
    }
  }
  F am = F(0), bm = F(0);
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (k == m) { am = a[k]; bm = b[k]; }
  F aa = fabs(am), ba = fabs(bm);
  F scl = s_isnan(aa) ? ba : (s_isnan(ba) ? aa : (aa > ba ? aa : ba));
  F inv = F(1) / scl;
  F s1[K], s2[K], t[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    F av = F(0), bv = F(0);
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (j == k + m) { av = a[j]; bv = b[j]; }
    s1[k] = (k + m < K) ? av * inv : F(0);
    s2[k] = (k + m < K) ? bv * inv : F(0);
  }
  mul<F, K>(s1, s1, t);
  mul<F, K>(s2, s2, s1);
#pragma unroll
  for (int k = 0; k < K; ++k) t[k] = t[k] + s1[k];
  sqrt_<F, K>(t, s1);
#pragma unroll
  for (int k = 0; k < K; ++k) t[k] = s1[k] * scl;
  t[0] = hypot(am, bm);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    F v = F(0);
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (j + m == k) v = t[j];
    c[k] = (k < m) ? F(0) : v;
  }
}
template <class F, int K> EH_DEV void atan2_(const F* a, const F* b, F* c) {
  F s1[K];
  if (b[0] != F(0)) {
    div<F, K>(a, b, s1);
    atan_<F, K>(s1, c);
    c[0] = atan2(a[0], b[0]);
  } else if (a[0] != F(0)) {
    div<F, K>(b, a, s1);
    atan_<F, K>(s1, c);
#pragma unroll
    for (int k = 0; k < K; ++k) c[k] = -c[k];
    c[0] = atan2(a[0], b[0]);
  } else {
    c[0] = F(0);
#pragma unroll
    for (int k = 1; k < K; ++k) c[k] = F(0);
  }
}
}  // namespace jt

template <class F, int K> EH_DEV Jet<F, K> jc(F v) {
  Jet<F, K> r;
  r.c[0] = v;
#pragma unroll
  for (int k = 1; k < K; ++k) r.c[k] = F(0);
  return r;
}

// Jet value of one op (eval_forward at T = Taylor<F,K>).
template <class F, int K> EH_DEV Jet<F, K> j_eval(uint32_t op, const Jet<F, K>& a, const Jet<F, K>& b) {
  Jet<F, K> r;
  switch (op) {
    case OP_ADD:
#pragma unroll
      for (int k = 0; k < K; ++k) r.c[k] = a.c[k] + b.c[k];
      return r;
    case OP_SUB:
#pragma unroll
      for (int k = 0; k < K; ++k) r.c[k] = a.c[k] - b.c[k];
      return r;
    case OP_MUL: jt::mul<F, K>(a.c, b.c, r.c); return r;
    case OP_DIV: jt::div<F, K>(a.c, b.c, r.c); return r;
    case OP_REM: {
      if (b.c[0] == F(0)) { jt::nanj<F, K>(r.c); return r; }
      F q = trunc(a.c[0] / b.c[0]);
      r.c[0] = fmod(a.c[0], b.c[0]);
#pragma unroll
      for (int k = 1; k < K; ++k) r.c[k] = a.c[k] - b.c[k] * q;
      return r;
    }
    case OP_POWF: jt::powf_<F, K>(a.c, b.c, r.c); return r;
    case OP_ATAN2: jt::atan2_<F, K>(a.c, b.c, r.c); return r;
    case OP_HYPOT: jt::hypot_<F, K>(a.c, b.c, r.c); return r;
    case OP_MAX: return (a.c[0] >= b.c[0] || s_isnan(b.c[0])) ? a : b;
    case OP_MIN: return (a.c[0] <= b.c[0] || s_isnan(b.c[0])) ? a : b;
    case OP_NEG:
#pragma unroll
      for (int k = 0; k < K; ++k) r.c[k] = -a.c[k];
      return r;
    case OP_RECIP: jt::recip<F, K>(a.c, r.c); return r;
    case OP_SQRT: jt::sqrt_<F, K>(a.c, r.c); return r;
    case OP_CBRT: jt::cbrt_<F, K>(a.c, r.c); return r;
    case OP_EXP: jt::exp_<F, K>(a.c, r.c); return r;
    case OP_EXP2: {
      F s[K];
      F ln2 = log(F(2));
      jt::scale<F, K>(a.c, ln2, s);
      jt::exp_<F, K>(s, r.c);
      r.c[0] = exp2(a.c[0]);
      return r;
    }
    case OP_EXPM1: jt::exp_<F, K>(a.c, r.c); r.c[0] = expm1(a.c[0]); return r;
    case OP_LN: jt::ln_<F, K>(a.c, r.c); return r;
    case OP_LOG2: {
      jt::ln_<F, K>(a.c, r.c);
      F inv = F(1) / log(F(2));
      r.c[0] = log2(a.c[0]);
#pragma unroll
      for (int k = 1; k < K; ++k) r.c[k] = r.c[k] * inv;
      return r;
    }
    case OP_LOG10: {
      jt::ln_<F, K>(a.c, r.c);
      F inv = F(1) / log(F(10));
      r.c[0] = log10(a.c[0]);
#pragma unroll
      for (int k = 1; k < K; ++k) r.c[k] = r.c[k] * inv;
      return r;
    }
    case OP_LN1P: {
      F s[K];
#pragma unroll
      for (int k = 1; k < K; ++k) s[k] = a.c[k];
      s[0] = F(1) + a.c[0];
      jt::ln_<F, K>(s, r.c);
      r.c[0] = log1p(a.c[0]);
      return r;
    }
    case OP_SIN: { F co[K]; jt::sin_cos<F, K>(a.c, r.c, co); return r; }
    case OP_COS: { F s[K]; jt::sin_cos<F, K>(a.c, s, r.c); return r; }
    case OP_TAN: jt::tan_<F, K>(a.c, r.c); return r;
    case OP_ASIN: jt::asin_<F, K>(a.c, r.c); return r;
    case OP_ACOS:
      jt::asin_<F, K>(a.c, r.c);
      r.c[0] = acos(a.c[0]);
#pragma unroll
      for (int k = 1; k < K; ++k) r.c[k] = -r.c[k];
      return r;
    case OP_ATAN: jt::atan_<F, K>(a.c, r.c); return r;
    case OP_SINH: { F ch[K]; jt::sinh_cosh<F, K>(a.c, r.c, ch); return r; }
    case OP_COSH: { F sh[K]; jt::sinh_cosh<F, K>(a.c, sh, r.c); return r; }
    case OP_TANH: jt::tanh_<F, K>(a.c, r.c); return r;
    case OP_ASINH: jt::asinh_<F, K>(a.c, r.c); return r;
    case OP_ACOSH: jt::acosh_<F, K>(a.c, r.c); return r;
    case OP_ATANH: jt::atanh_<F, K>(a.c, r.c); return r;
    case OP_ABS: {
      F sign = F(1);
      bool found = a.c[0] != F(0);
      if (found) sign = s_signum(a.c[0]);
#pragma unroll
      for (int k = 1; k < K; ++k)
        if (!found && a.c[k] != F(0)) { sign = s_signum(a.c[k]); found = true; }
#pragma unroll
      for (int k = 0; k < K; ++k) r.c[k] = a.c[k] * sign;
      return r;
    }
    case OP_SIGNUM: return jc<F, K>(s_signum(a.c[0]));
    case OP_FLOOR: return jc<F, K>(floor(a.c[0]));
    case OP_CEIL: return jc<F, K>(ceil(a.c[0]));
    case OP_ROUND: return jc<F, K>(round(a.c[0]));
    case OP_TRUNC: return jc<F, K>(trunc(a.c[0]));
    case OP_FRACT: r = a; r.c[0] = s_fract(a.c[0]); return r;
    default: jt::nanj<F, K>(r.c); return r;
  }
}
template <class F, int K> EH_DEV Jet<F, K> j_powi(const Jet<F, K>& a, int32_t n) {
  Jet<F, K> r;
  jt::powi_<F, K>(a.c, n, r.c);
  return r;
}

}  // namespace ehip
