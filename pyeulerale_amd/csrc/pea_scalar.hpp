// Scalar types the discretization is instantiated on.
//
//   double   -- the FP64 residual and the adjoint (VJP) kernels;
//   Cx       -- complex128 for the complex-step path (the reference's own method, whose Jacobians
//               are Im R(u + i h e)/h, spatial_discretization.f90:45-53);
//   Dual<K>  -- a value with K tangent lanes for the hand-written forward mode (JVP), and for
//               tangent-over-adjoint second-order products when the adjoint runs on it.
//
// Every type provides the same operations, so one templated body serves all of them:
//   +, -, *, / (with each other and with double), ssqrt, sabs, val (the real/primal part).
// sabs carries the complex-step convention d|x| = sign(1, x) dx with a signed zero
// (abs_c, f90:97-102); branch decisions (the biased maximum) read val() only.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define PEA_HD __host__ __device__ inline __attribute__((always_inline))

namespace pea {

constexpr double kGamma = 1.4;                                   // f90:55
constexpr double kDegToRad = 0.0174532925199432957692369;        // f90:54
constexpr double kMaxBias = 1e-8;                                // f90:115
constexpr double kCsStep = 1e-12;                                // f90:52

// ------------------------------------------------------------------------------- double
PEA_HD double val(double x) { return x; }
PEA_HD double sabs(double x) { return fabs(x); }
PEA_HD double ssqrt(double x) { return sqrt(x); }
PEA_HD double sign1(double x) { return copysign(1.0, x); }

// ------------------------------------------------------------------------------- complex
struct Cx {
    double r, i;
};
// Complex-step arithmetic is never contracted into FMAs: the reference evaluates R in complex
// arithmetic with separately rounded products and sums, and at kinks of abs/max (e.g. the
// pressure sensor of a wall face at the free stream, where p_ll - 2 p_l + p_r is round-off)
// the sign of a round-off-sized real part selects the derivative branch.  Matching the
// reference's rounding keeps its Jacobians, and so its Newton iterates, reproducible.
#define PEA_NOFMA _Pragma("clang fp contract(off)")
PEA_HD Cx operator+(Cx a, Cx b) { PEA_NOFMA return {a.r + b.r, a.i + b.i}; }
PEA_HD Cx operator-(Cx a, Cx b) { PEA_NOFMA return {a.r - b.r, a.i - b.i}; }
PEA_HD Cx operator-(Cx a) { return {-a.r, -a.i}; }
PEA_HD Cx operator*(Cx a, Cx b) {
    PEA_NOFMA return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
PEA_HD Cx operator*(double s, Cx a) { PEA_NOFMA return {s * a.r, s * a.i}; }
PEA_HD Cx operator*(Cx a, double s) { PEA_NOFMA return {a.r * s, a.i * s}; }
PEA_HD Cx operator/(Cx a, double s) { PEA_NOFMA return {a.r / s, a.i / s}; }
PEA_HD Cx operator+(Cx a, double s) { PEA_NOFMA return {a.r + s, a.i}; }
PEA_HD Cx operator-(Cx a, double s) { PEA_NOFMA return {a.r - s, a.i}; }
PEA_HD Cx operator-(double s, Cx a) { PEA_NOFMA return {s - a.r, -a.i}; }
PEA_HD Cx operator+(double s, Cx a) { PEA_NOFMA return {s + a.r, a.i}; }
// Smith's scaled division (the range reduction Fortran compilers use for complex division).
PEA_HD Cx operator/(Cx a, Cx b) {
    PEA_NOFMA
    if (fabs(b.r) >= fabs(b.i)) {
        double t = b.i / b.r, d = b.r + b.i * t;
        return {(a.r + a.i * t) / d, (a.i - a.r * t) / d};
    }
    double t = b.r / b.i, d = b.r * t + b.i;
    return {(a.r * t + a.i) / d, (a.i * t - a.r) / d};
}
PEA_HD Cx operator/(double s, Cx b) { return Cx{s, 0.0} / b; }
PEA_HD double val(Cx x) { return x.r; }
PEA_HD Cx sabs(Cx x) { return {fabs(x.r), x.i * copysign(1.0, x.r)}; }
// principal square root, computed without cancellation
PEA_HD Cx ssqrt(Cx z) {
    PEA_NOFMA
    if (z.r == 0.0 && z.i == 0.0) return {0.0, z.i};
    double t = sqrt(0.5 * (fabs(z.r) + hypot(z.r, z.i)));
    if (z.r >= 0.0) return {t, z.i / (2.0 * t)};
    return {fabs(z.i) / (2.0 * t), copysign(t, z.i)};
}

// Reciprocal for derivative factors: the hardware estimate refined by two Newton steps
// (r <- r + r(1 - x r)), within a few ulp of 1/x for normal x.  Primal values keep IEEE division
// so that every branch decision matches the residual; tangent and adjoint factors only need
// relative accuracy far below the 1e-10 product tolerance.
PEA_HD double rcp_fast(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;
#endif
}

// sqrt(x) for a positive normal x (an edge length from its squared components): the hardware
// reciprocal square root estimate refined by one Goldschmidt step and one Newton correction,
// within about 1 ulp of the correctly rounded root, without the range scaling and special-value
// handling of IEEE sqrt.  x == 0 gives 0; a negative or NaN x gives NaN.
PEA_HD double sqrt_pos(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    const double r = __builtin_amdgcn_rsq(x);
    double s = x * r, h = 0.5 * r;
    const double e = fma(-h, s, 0.5);
    s = fma(s, e, s);
    h = fma(h, e, h);
    const double d = fma(-s, s, x);
    s = fma(d, h, s);
    return x == 0.0 ? 0.0 : s;
#else
    return sqrt(x);
#endif
}

// ------------------------------------------------------------------------------- dual
template <int K>
struct Dual {
    double v;
    double d[K];
};
template <int K>
PEA_HD Dual<K> dconst(double v) {
    Dual<K> r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = 0.0;
    return r;
}
template <int K>
PEA_HD Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int K>
PEA_HD Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int K>
PEA_HD Dual<K> operator-(const Dual<K>& a) {
    Dual<K> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int K>
PEA_HD Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int K>
PEA_HD Dual<K> operator*(double s, const Dual<K>& a) {
    Dual<K> r;
    r.v = s * a.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = s * a.d[k];
    return r;
}
template <int K>
PEA_HD Dual<K> operator*(const Dual<K>& a, double s) {
    return s * a;
}
template <int K>
PEA_HD Dual<K> operator/(const Dual<K>& a, double s) {
    Dual<K> r;
    r.v = a.v / s;
    // a literal divisor (the halves of the central and ALE terms) folds to an exact constant
    const double inv = __builtin_constant_p(s) ? 1.0 / s : rcp_fast(s);
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * inv;
    return r;
}
template <int K>
PEA_HD Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v / b.v;
    const double inv = rcp_fast(b.v);
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * inv;
    return r;
}
template <int K>
PEA_HD Dual<K> operator+(const Dual<K>& a, double s) {
    Dual<K> r = a;
    r.v = a.v + s;
    return r;
}
template <int K>
PEA_HD Dual<K> operator-(const Dual<K>& a, double s) {
    Dual<K> r = a;
    r.v = a.v - s;
    return r;
}
template <int K>
PEA_HD Dual<K> operator-(double s, const Dual<K>& a) {
    Dual<K> r = -a;
    r.v = s - a.v;
    return r;
}
template <int K>
PEA_HD Dual<K> operator/(double s, const Dual<K>& b) {
    Dual<K> r;
    r.v = s / b.v;
    const double g = -r.v * rcp_fast(b.v);
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = g * b.d[k];
    return r;
}
template <int K>
PEA_HD Dual<K> operator+(double s, const Dual<K>& a) {
    return a + s;
}
template <int K>
PEA_HD double val(const Dual<K>& x) {
    return x.v;
}
template <int K>
PEA_HD Dual<K> sabs(const Dual<K>& x) {
    const double s = copysign(1.0, x.v);
    Dual<K> r;
    r.v = fabs(x.v);
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = s * x.d[k];
    return r;
}
template <int K>
PEA_HD Dual<K> ssqrt(const Dual<K>& x) {
    Dual<K> r;
    r.v = sqrt(x.v);
    const double h = 0.5 * rcp_fast(r.v);
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = x.d[k] * h;
    return r;
}
template <int K>
PEA_HD double sign1(const Dual<K>& x) {
    return copysign(1.0, x.v);
}

// a/b through rcp_fast, within about 1.5 ulp of the correctly rounded quotient: the face-based
// kernels' divisions (the reference's expressions, one rounding apart from its IEEE divisions;
// R agrees to a few ulp, far inside the 1e-12 parity tolerance).
PEA_HD double qdiv(double a, double b) { return a * rcp_fast(b); }
template <int K>
PEA_HD Dual<K> qdiv(const Dual<K>& a, const Dual<K>& b) {
    const double rb = rcp_fast(b.v);
    Dual<K> r;
    r.v = a.v * rb;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * rb;
    return r;
}
template <int K>
PEA_HD Dual<K> qdiv(const Dual<K>& a, double b) {
    const double rb = rcp_fast(b);
    Dual<K> r;
    r.v = a.v * rb;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * rb;
    return r;
}

// 1/x where it only scales derivatives (adjoint factors): rcp_fast for double; for Dual<K>
// the quotient's own rule.
PEA_HD double recip(double x) { return rcp_fast(x); }
template <int K>
PEA_HD Dual<K> recip(const Dual<K>& x) {
    return 1.0 / x;
}

// Constant of scalar type T.
template <class T>
struct Lit {
    PEA_HD static T make(double v) { return T(v); }
};
template <>
struct Lit<Cx> {
    PEA_HD static Cx make(double v) { return {v, 0.0}; }
};
template <int K>
struct Lit<Dual<K>> {
    PEA_HD static Dual<K> make(double v) { return dconst<K>(v); }
};
template <class T>
PEA_HD T lit(double v) {
    return Lit<T>::make(v);
}

// Start value of an adjoint sum: -0.0, the identity of IEEE addition (-0 + x == x for every x,
// unlike +0 + (-0)), so a sum that starts from it folds its first addition away.
template <class T>
PEA_HD T adj_zero() {
    return lit<T>(-0.0);
}

// Biased maximum: x only if val(x) exceeds val(y) by more than 1e-8, else y (f90:111-120).
// The bias makes the choice depend on argument order; callers keep the reference's order.
template <class T>
PEA_HD bool maxb_takes_first(const T& x, const T& y) {
    return val(x) > val(y) + kMaxBias;
}
template <class T>
PEA_HD T maxb(const T& x, const T& y) {
    return maxb_takes_first(x, y) ? x : y;
}

}  // namespace pea
