// Scalar types the kernel bodies are instantiated with besides float and
// double:
//  - Dual<T>: value + one derivative (forward-mode differentiation), used to
//    build exact local Jacobians of the step's one-shot kernels for the
//    discrete adjoint;
//  - CStep: a complex number in double (complex-step differentiation,
//    f'(x) v = Im f(x + i h v) / h), used for the FP64 reference gradient.
// Comparisons and branches use the value / real part, so the branch taken is
// the one of the undifferentiated run.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace uws {
// the types and their math live in a nested namespace so that argument-
// dependent lookup finds them while float/double math keeps its overloads
namespace sc {

#define UWS_HD __host__ __device__ __forceinline__

template <class T>
struct Dual {
    T v, d;
    UWS_HD Dual() : v(0), d(0) {}
    UWS_HD Dual(T a) : v(a), d(0) {}
    UWS_HD Dual(T a, T b) : v(a), d(b) {}
    template <class U, class = typename std::enable_if<std::is_arithmetic<U>::value>::type>
    UWS_HD Dual(U a) : v((T)a), d(0) {}
    UWS_HD Dual& operator+=(const Dual& o) { v += o.v; d += o.d; return *this; }
    UWS_HD Dual& operator-=(const Dual& o) { v -= o.v; d -= o.d; return *this; }
    UWS_HD Dual& operator*=(const Dual& o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
    UWS_HD Dual& operator/=(const Dual& o) { *this = *this / o; return *this; }
    UWS_HD Dual operator-() const { return Dual(-v, -d); }
    UWS_HD friend Dual operator+(Dual a, const Dual& b) { return a += b; }
    UWS_HD friend Dual operator-(Dual a, const Dual& b) { return a -= b; }
    UWS_HD friend Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
    UWS_HD friend Dual operator/(const Dual& a, const Dual& b) {
        const T q = a.v / b.v;
        return Dual(q, (a.d - q * b.d) / b.v);
    }
    UWS_HD friend bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
    UWS_HD friend bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
    UWS_HD friend bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
    UWS_HD friend bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
    UWS_HD explicit operator double() const { return (double)v; }
    UWS_HD explicit operator float() const { return (float)v; }
};
template <class T>
UWS_HD Dual<T> sqrt(const Dual<T>& a) {
    const T s = ::sqrt(a.v);
    return Dual<T>(s, s > (T)0 ? a.d / ((T)2 * s) : (T)0);
}
template <class T>
UWS_HD Dual<T> log(const Dual<T>& a) { return Dual<T>(::log(a.v), a.d / a.v); }
template <class T>
UWS_HD Dual<T> fabs(const Dual<T>& a) { return a.v < (T)0 ? -a : a; }
template <class T>
UWS_HD bool isfinite(const Dual<T>& a) { return ::isfinite(a.v) && ::isfinite(a.d); }

struct CStep {
    double re, im;
    UWS_HD CStep() : re(0), im(0) {}
    UWS_HD CStep(double a) : re(a), im(0) {}
    UWS_HD CStep(double a, double b) : re(a), im(b) {}
    template <class U, class = typename std::enable_if<std::is_arithmetic<U>::value>::type>
    UWS_HD CStep(U a) : re((double)a), im(0) {}
    UWS_HD CStep& operator+=(const CStep& o) { re += o.re; im += o.im; return *this; }
    UWS_HD CStep& operator-=(const CStep& o) { re -= o.re; im -= o.im; return *this; }
    UWS_HD CStep& operator*=(const CStep& o) { *this = *this * o; return *this; }
    UWS_HD CStep& operator/=(const CStep& o) { *this = *this / o; return *this; }
    UWS_HD CStep operator-() const { return CStep(-re, -im); }
    UWS_HD friend CStep operator+(CStep a, const CStep& b) { return a += b; }
    UWS_HD friend CStep operator-(CStep a, const CStep& b) { return a -= b; }
    UWS_HD friend CStep operator*(const CStep& a, const CStep& b) {
        return CStep(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
    }
    UWS_HD friend CStep operator/(const CStep& a, const CStep& b) {
        const double den = b.re * b.re + b.im * b.im;
        return CStep((a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den);
    }
    UWS_HD friend bool operator<(const CStep& a, const CStep& b) { return a.re < b.re; }
    UWS_HD friend bool operator>(const CStep& a, const CStep& b) { return a.re > b.re; }
    UWS_HD friend bool operator<=(const CStep& a, const CStep& b) { return a.re <= b.re; }
    UWS_HD friend bool operator>=(const CStep& a, const CStep& b) { return a.re >= b.re; }
    UWS_HD explicit operator double() const { return re; }
    UWS_HD explicit operator float() const { return (float)re; }
};
// principal branches; arguments here have positive real parts
UWS_HD CStep sqrt(const CStep& a) {
    // at a zero real part the derivative is taken as 0, the convention of Dual
    if (a.re == 0.0) return CStep(0.0, 0.0);
    const double m = ::sqrt(a.re * a.re + a.im * a.im);
    const double r = ::sqrt(0.5 * (m + a.re));
    const double i = r > 0.0 ? a.im / (2.0 * r) : ::sqrt(0.5 * (m - a.re));
    return CStep(r, i);
}
UWS_HD CStep log(const CStep& a) { return CStep(::log(::sqrt(a.re * a.re + a.im * a.im)), ::atan2(a.im, a.re)); }
// analytic continuation of |x| (the branch of the real part)
UWS_HD CStep fabs(const CStep& a) { return a.re < 0.0 ? -a : a; }
UWS_HD bool isfinite(const CStep& a) { return ::isfinite(a.re) && ::isfinite(a.im); }

// value part of any scalar
UWS_HD double real_of(float x) { return x; }
UWS_HD double real_of(double x) { return x; }
template <class T>
UWS_HD double real_of(const Dual<T>& x) { return (double)x.v; }
UWS_HD double real_of(const CStep& x) { return x.re; }

}  // namespace sc
using sc::CStep;
using sc::Dual;
using sc::real_of;

}  // namespace uws
