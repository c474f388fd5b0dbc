/* Complex.h -- interleaved FP64 complex number with the member layout and the
 * operator surface of the reference's Complex class (Complex.h:30-296):
 * public `re`, `im`, arithmetic with Complex and double, conj, abs2, abs,
 * expi.  Layout-compatible with `double _Complex` and numpy complex128.
 */
#ifndef FFTWPP_AMD_COMPLEX_H
#define FFTWPP_AMD_COMPLEX_H

#include <cmath>
#include <iostream>

class Complex {
public:
  double re;
  double im;

  Complex() : re(0.0), im(0.0) {}
  Complex(double r) : re(r), im(0.0) {}
  Complex(double r, double i) : re(r), im(i) {}

  double real() const { return re; }
  double imag() const { return im; }

  Complex &operator+=(const Complex &b) { re += b.re; im += b.im; return *this; }
  Complex &operator-=(const Complex &b) { re -= b.re; im -= b.im; return *this; }
  Complex &operator*=(const Complex &b) {
    double r=re*b.re-im*b.im;
    im=re*b.im+im*b.re;
    re=r;
    return *this;
  }
  Complex &operator*=(double s) { re *= s; im *= s; return *this; }
  Complex &operator/=(double s) { re /= s; im /= s; return *this; }
  Complex &operator/=(const Complex &b) {
    double d=b.re*b.re+b.im*b.im;
    double r=(re*b.re+im*b.im)/d;
    im=(im*b.re-re*b.im)/d;
    re=r;
    return *this;
  }
};

inline Complex operator+(Complex a, const Complex &b) { return a += b; }
inline Complex operator-(Complex a, const Complex &b) { return a -= b; }
inline Complex operator*(Complex a, const Complex &b) { return a *= b; }
inline Complex operator/(Complex a, const Complex &b) { return a /= b; }
inline Complex operator*(Complex a, double s) { return a *= s; }
inline Complex operator*(double s, Complex a) { return a *= s; }
inline Complex operator/(Complex a, double s) { return a /= s; }
inline Complex operator+(Complex a, double s) { a.re += s; return a; }
inline Complex operator+(double s, Complex a) { a.re += s; return a; }
inline Complex operator-(Complex a, double s) { a.re -= s; return a; }
inline Complex operator-(double s, const Complex &a) { return Complex(s-a.re,-a.im); }
inline Complex operator-(const Complex &a) { return Complex(-a.re,-a.im); }
inline bool operator==(const Complex &a, const Complex &b) { return a.re == b.re && a.im == b.im; }
inline bool operator!=(const Complex &a, const Complex &b) { return !(a == b); }

inline Complex conj(const Complex &a) { return Complex(a.re,-a.im); }
inline double real(const Complex &a) { return a.re; }
inline double imag(const Complex &a) { return a.im; }
inline double abs2(const Complex &a) { return a.re*a.re+a.im*a.im; }
inline double abs(const Complex &a) { return std::hypot(a.re,a.im); }
inline double arg(const Complex &a) { return std::atan2(a.im,a.re); }
// e^{i theta}
inline Complex expi(double theta) { return Complex(std::cos(theta),std::sin(theta)); }

inline std::ostream &operator<<(std::ostream &s, const Complex &a)
{
  return s << "(" << a.re << "," << a.im << ")";
}

#endif
