// ORACLE — test infrastructure only. This header belongs to the FP64 CPU
// restatement of the reference's constrained-HMC path. Only tests/, the
// smoke check and the benchmark's CPU-baseline leg may use it; the shipped
// library (mag_amd/) never includes or links anything under oracle/.
//
// 3x3 complex FP64 matrix with the operations the reference takes from
// cps::Matrix (CPS, un-vendored). Layout follows the reference's implied
// cps::Matrix layout: element (i,j) at reals 2(3i+j) (re) and 2(3i+j)+1 (im),
// row-major (alg_gchb.h:59-85; 12-real memcpy at field-matrix.h:92-95).
// Semantics are restated from the textbook definitions:
//   Dagger, ReTr, Tr, UnitMatrix, ZeroMatrix,
//   TrLessAntiHermMatrix: (M - M^+)/2 - Tr(M - M^+)/6
//   Unitarize: Gram-Schmidt on rows 0,1; row2 = conj(row0 x row1).
// Complex products are written out (a*c - b*d, a*d + b*c) so that results
// do not depend on libstdc++'s NaN-aware complex multiply.
#pragma once
#include <cmath>

namespace orc {

struct cplx {
  double re, im;
  cplx() : re(0.), im(0.) {}
  cplx(double r, double i = 0.) : re(r), im(i) {}
};
inline cplx operator+(cplx a, cplx b) { return cplx(a.re + b.re, a.im + b.im); }
inline cplx operator-(cplx a, cplx b) { return cplx(a.re - b.re, a.im - b.im); }
inline cplx operator-(cplx a) { return cplx(-a.re, -a.im); }
inline cplx operator*(cplx a, cplx b) {
  return cplx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
inline cplx operator*(double s, cplx a) { return cplx(s * a.re, s * a.im); }
inline cplx operator*(cplx a, double s) { return cplx(s * a.re, s * a.im); }
inline cplx operator/(cplx a, double s) { return cplx(a.re / s, a.im / s); }
inline cplx operator/(cplx a, cplx b) {
  double d = b.re * b.re + b.im * b.im;
  return cplx((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
inline cplx conj(cplx a) { return cplx(a.re, -a.im); }
inline double norm2(cplx a) { return a.re * a.re + a.im * a.im; }
inline cplx expix(double x) { return cplx(std::cos(x), std::sin(x)); }  // field-matrix.h:67-69
inline cplx I() { return cplx(0., 1.); }                                  // field-matrix.h:71-73

struct Mat {
  cplx e[9];
  cplx& operator()(int i, int j) { return e[3 * i + j]; }
  const cplx& operator()(int i, int j) const { return e[3 * i + j]; }
  static Mat zero() { return Mat(); }
  static Mat unit() {
    Mat m;
    m.e[0] = m.e[4] = m.e[8] = cplx(1.);
    return m;
  }
  void load(const double* p) {
    for (int k = 0; k < 9; ++k) e[k] = cplx(p[2 * k], p[2 * k + 1]);
  }
  void store(double* p) const {
    for (int k = 0; k < 9; ++k) { p[2 * k] = e[k].re; p[2 * k + 1] = e[k].im; }
  }
  Mat& operator+=(const Mat& b) { for (int k = 0; k < 9; ++k) e[k] = e[k] + b.e[k]; return *this; }
  Mat& operator-=(const Mat& b) { for (int k = 0; k < 9; ++k) e[k] = e[k] - b.e[k]; return *this; }
};

inline Mat operator*(const Mat& a, const Mat& b) {
  Mat c;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cplx s = a(i, 0) * b(0, j);
      s = s + a(i, 1) * b(1, j);
      s = s + a(i, 2) * b(2, j);
      c(i, j) = s;
    }
  return c;
}
inline Mat operator+(const Mat& a, const Mat& b) { Mat c = a; c += b; return c; }
inline Mat operator-(const Mat& a, const Mat& b) { Mat c = a; c -= b; return c; }
inline Mat operator*(const Mat& a, cplx s) { Mat c; for (int k = 0; k < 9; ++k) c.e[k] = a.e[k] * s; return c; }
inline Mat operator*(const Mat& a, double s) { Mat c; for (int k = 0; k < 9; ++k) c.e[k] = a.e[k] * s; return c; }
inline Mat dagger(const Mat& a) {
  Mat c;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c(i, j) = conj(a(j, i));
  return c;
}
inline cplx tr(const Mat& a) { return a.e[0] + a.e[4] + a.e[8]; }
inline double retr(const Mat& a) { return a.e[0].re + a.e[4].re + a.e[8].re; }
// Re Tr(A B) without forming the product.
inline double retr_mul(const Mat& a, const Mat& b) {
  double s = 0.;
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) s += a(i, k).re * b(k, i).re - a(i, k).im * b(k, i).im;
  return s;
}
// (M - M^+)/2 - Tr(M - M^+)/6 * 1
inline Mat trless_antiherm(const Mat& m) {
  Mat a;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a(i, j) = (m(i, j) - conj(m(j, i))) * 0.5;
  cplx t = tr(a) / 3.;
  for (int i = 0; i < 3; ++i) a(i, i) = a(i, i) - t;
  return a;
}
// ((M + M^+) - Tr(M + M^+)/3) / 2, field-multi.h:41-45
inline Mat hermitian_traceless(const Mat& m) {
  Mat s = m + dagger(m);
  cplx t = tr(s) / 3.;
  Mat r = s;
  for (int i = 0; i < 3; ++i) r(i, i) = r(i, i) - t;
  return r * 0.5;
}
inline double frob(const Mat& m) {  // field-matrix.h:75-81
  double s = 0.;
  for (int k = 0; k < 9; ++k) s += norm2(m.e[k]);
  return std::sqrt(s);
}
inline cplx det(const Mat& m) {
  return m(0, 0) * (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1)) -
         m(0, 1) * (m(1, 0) * m(2, 2) - m(1, 2) * m(2, 0)) +
         m(0, 2) * (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0));
}
// Project onto SU(3): normalise row 0, orthogonalise and normalise row 1,
// row 2 = conj(row0 x row1).
inline void unitarize(Mat& m) {
  double n0 = 0.;
  for (int j = 0; j < 3; ++j) n0 += norm2(m(0, j));
  n0 = 1. / std::sqrt(n0);
  for (int j = 0; j < 3; ++j) m(0, j) = m(0, j) * n0;
  cplx c;
  for (int j = 0; j < 3; ++j) c = c + conj(m(0, j)) * m(1, j);
  for (int j = 0; j < 3; ++j) m(1, j) = m(1, j) - c * m(0, j);
  double n1 = 0.;
  for (int j = 0; j < 3; ++j) n1 += norm2(m(1, j));
  n1 = 1. / std::sqrt(n1);
  for (int j = 0; j < 3; ++j) m(1, j) = m(1, j) * n1;
  m(2, 0) = conj(m(0, 1) * m(1, 2) - m(0, 2) * m(1, 1));
  m(2, 1) = conj(m(0, 2) * m(1, 0) - m(0, 0) * m(1, 2));
  m(2, 2) = conj(m(0, 0) * m(1, 1) - m(0, 1) * m(1, 0));
}

}  // namespace orc
