// Device-side SU(3) algebra for gfx950 (also compiled for the host so the
// C-ABI test hooks can share it). 3x3 complex matrices live in registers as
// 9 (re, im) pairs; all products are written with fused multiply-adds on the
// vector ALUs (no MFMA: a 3x3 complex product is not a dense contraction
// worth a matrix-core tile).
//
// Conventions (restated from the reference, hummingtree/mag):
//   element (i,j) <-> reals 2(3i+j), 2(3i+j)+1            (alg_gchb.h:59-85)
//   su(3) generators T_a = lambda_a / sqrt(2)              (field-hmc.h:30-71)
//   TA(M) = (M - M^+)/2 - Tr(M - M^+)/6                    (CPS TrLessAntiHermMatrix)
//   HT(M) = ((M + M^+) - Tr(M + M^+)/3) / 2                (field-multi.h:41-45)
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#define MAG_HD __host__ __device__ __forceinline__

namespace mag {

template <class T>
struct cpx {
  T r, i;
};

MAG_HD float mfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
MAG_HD double mfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class T>
MAG_HD cpx<T> mk(T r, T i) { return cpx<T>{r, i}; }
template <class T>
MAG_HD cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.r + b.r, a.i + b.i}; }
template <class T>
MAG_HD cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.r - b.r, a.i - b.i}; }
template <class T>
MAG_HD cpx<T> operator*(cpx<T> a, cpx<T> b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
template <class T>
MAG_HD cpx<T> operator*(T s, cpx<T> a) { return {s * a.r, s * a.i}; }
template <class T>
MAG_HD cpx<T> cconj(cpx<T> a) { return {a.r, -a.i}; }
// a + b*c
template <class T>
MAG_HD cpx<T> cfma(cpx<T> b, cpx<T> c, cpx<T> a) {
  return {mfma(b.r, c.r, mfma(-b.i, c.i, a.r)), mfma(b.r, c.i, mfma(b.i, c.r, a.i))};
}
// a + conj(b)*c
template <class T>
MAG_HD cpx<T> cfma_cj(cpx<T> b, cpx<T> c, cpx<T> a) {
  return {mfma(b.r, c.r, mfma(b.i, c.i, a.r)), mfma(b.r, c.i, mfma(-b.i, c.r, a.i))};
}

template <class T>
struct M3 {
  cpx<T> e[9];
  MAG_HD cpx<T>& operator()(int i, int j) { return e[3 * i + j]; }
  MAG_HD const cpx<T>& operator()(int i, int j) const { return e[3 * i + j]; }
};

template <class T>
MAG_HD M3<T> zero3() {
  M3<T> m;
#pragma unroll
  for (int k = 0; k < 9; ++k) m.e[k] = {T(0), T(0)};
  return m;
}
template <class T>
MAG_HD M3<T> unit3() {
  M3<T> m = zero3<T>();
  m.e[0].r = m.e[4].r = m.e[8].r = T(1);
  return m;
}
template <class TO, class TI>
MAG_HD M3<TO> convert(const M3<TI>& a) {
  M3<TO> m;
#pragma unroll
  for (int k = 0; k < 9; ++k) m.e[k] = {TO(a.e[k].r), TO(a.e[k].i)};
  return m;
}

// C = A B
template <class T>
MAG_HD M3<T> mul(const M3<T>& a, const M3<T>& b) {
  M3<T> c;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = a(i, 0) * b(0, j);
      s = cfma(a(i, 1), b(1, j), s);
      s = cfma(a(i, 2), b(2, j), s);
      c(i, j) = s;
    }
  return c;
}
// C = A B^+
template <class T>
MAG_HD M3<T> mul_nd(const M3<T>& a, const M3<T>& b) {
  M3<T> c;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = a(i, 0) * cconj(b(j, 0));
      s = cfma(a(i, 1), cconj(b(j, 1)), s);
      s = cfma(a(i, 2), cconj(b(j, 2)), s);
      c(i, j) = s;
    }
  return c;
}
// C = A^+ B
template <class T>
MAG_HD M3<T> mul_dn(const M3<T>& a, const M3<T>& b) {
  M3<T> c;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = cconj(a(0, i)) * b(0, j);
      s = cfma_cj(a(1, i), b(1, j), s);
      s = cfma_cj(a(2, i), b(2, j), s);
      c(i, j) = s;
    }
  return c;
}
// Products with a Hermitian factor H (Q, Lambda): the diagonal of H is real,
// so three of the nine complex multiply-adds of every column (row) need only
// half their instructions. Same terms in the same order as the general
// products; the dropped ones are exact zeros.
template <class T>
MAG_HD cpx<T> rfma(T r, cpx<T> c, cpx<T> acc) { return {mfma(r, c.r, acc.r), mfma(r, c.i, acc.i)}; }
template <class T>
MAG_HD cpx<T> neg(cpx<T> a) { return {-a.r, -a.i}; }
// C = H A
template <class T>
MAG_HD M3<T> mul_hn(const M3<T>& h, const M3<T>& a) {
  M3<T> c;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = {T(0), T(0)};
#pragma unroll
      for (int k = 0; k < 3; ++k) s = (k == i) ? rfma(h(i, i).r, a(k, j), s) : cfma(h(i, k), a(k, j), s);
      c(i, j) = s;
    }
  return c;
}
// H^2 of a Hermitian H from its upper triangle and real diagonal: the
// diagonal is sum_k |h_ik|^2, the lower triangle the conjugate of the upper
// (a third of the multiply-adds of a general product).
template <class T>
MAG_HD M3<T> sq_h(const M3<T>& h) {
  const T d0 = h(0, 0).r, d1 = h(1, 1).r, d2 = h(2, 2).r;
  const cpx<T> u01 = h(0, 1), u02 = h(0, 2), u12 = h(1, 2);
  const T n01 = mfma(u01.r, u01.r, u01.i * u01.i);
  const T n02 = mfma(u02.r, u02.r, u02.i * u02.i);
  const T n12 = mfma(u12.r, u12.r, u12.i * u12.i);
  M3<T> r;
  r(0, 0) = {mfma(d0, d0, n01 + n02), T(0)};
  r(1, 1) = {mfma(d1, d1, n01 + n12), T(0)};
  r(2, 2) = {mfma(d2, d2, n02 + n12), T(0)};
  r(0, 1) = rfma(d0 + d1, u01, u02 * cconj(u12));
  r(0, 2) = rfma(d0 + d2, u02, u01 * u12);
  r(1, 2) = rfma(d1 + d2, u12, cconj(u01) * u02);
  r(1, 0) = cconj(r(0, 1));
  r(2, 0) = cconj(r(0, 2));
  r(2, 1) = cconj(r(1, 2));
  return r;
}
// Tr(A H) for a Hermitian H (real diagonal)
template <class T>
MAG_HD cpx<T> trace_mul_nh(const M3<T>& a, const M3<T>& h) {
  cpx<T> s = {T(0), T(0)};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) s = (k == i) ? rfma(h(i, i).r, a(i, i), s) : cfma(a(i, k), h(k, i), s);
  return s;
}
// C += A H
template <class T>
MAG_HD void mul_acc_nh(M3<T>& c, const M3<T>& a, const M3<T>& h) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = c(i, j);
#pragma unroll
      for (int k = 0; k < 3; ++k) s = (k == j) ? rfma(h(j, j).r, a(i, k), s) : cfma(a(i, k), h(k, j), s);
      c(i, j) = s;
    }
}
// C -= A H
template <class T>
MAG_HD void mul_sub_nh(M3<T>& c, const M3<T>& a, const M3<T>& h) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = c(i, j);
#pragma unroll
      for (int k = 0; k < 3; ++k) s = (k == j) ? rfma(-h(j, j).r, a(i, k), s) : cfma(neg(a(i, k)), h(k, j), s);
      c(i, j) = s;
    }
}
// C = A^+ H
template <class T>
MAG_HD M3<T> mul_dh(const M3<T>& a, const M3<T>& h) {
  M3<T> c;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = {T(0), T(0)};
#pragma unroll
      for (int k = 0; k < 3; ++k) s = (k == j) ? rfma(h(j, j).r, cconj(a(k, i)), s) : cfma_cj(a(k, i), h(k, j), s);
      c(i, j) = s;
    }
  return c;
}
// C -= H A^+
template <class T>
MAG_HD void mul_hd_sub(M3<T>& c, const M3<T>& h, const M3<T>& a) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = c(i, j);
#pragma unroll
      for (int k = 0; k < 3; ++k)
        s = (k == i) ? rfma(-h(i, i).r, cconj(a(j, k)), s) : cfma(neg(h(i, k)), cconj(a(j, k)), s);
      c(i, j) = s;
    }
}

// C = A^+ B^+ = (B A)^+
template <class T>
MAG_HD M3<T> mul_dd(const M3<T>& a, const M3<T>& b) {
  M3<T> c;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = cconj(a(0, i)) * cconj(b(j, 0));
      s = s + cconj(a(1, i)) * cconj(b(j, 1));
      s = s + cconj(a(2, i)) * cconj(b(j, 2));
      c(i, j) = s;
    }
  return c;
}
// C += A B
template <class T>
MAG_HD void mul_acc(M3<T>& c, const M3<T>& a, const M3<T>& b) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = c(i, j);
      s = cfma(a(i, 0), b(0, j), s);
      s = cfma(a(i, 1), b(1, j), s);
      s = cfma(a(i, 2), b(2, j), s);
      c(i, j) = s;
    }
}
// C += A B^+
template <class T>
MAG_HD void mul_nd_acc(M3<T>& c, const M3<T>& a, const M3<T>& b) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = c(i, j);
      s = cfma(a(i, 0), cconj(b(j, 0)), s);
      s = cfma(a(i, 1), cconj(b(j, 1)), s);
      s = cfma(a(i, 2), cconj(b(j, 2)), s);
      c(i, j) = s;
    }
}
// C += A^+ B
template <class T>
MAG_HD void mul_dn_acc(M3<T>& c, const M3<T>& a, const M3<T>& b) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = c(i, j);
      s = cfma_cj(a(0, i), b(0, j), s);
      s = cfma_cj(a(1, i), b(1, j), s);
      s = cfma_cj(a(2, i), b(2, j), s);
      c(i, j) = s;
    }
}

template <class T>
MAG_HD M3<T> dag(const M3<T>& a) {
  M3<T> c;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) c(i, j) = cconj(a(j, i));
  return c;
}
template <class T>
MAG_HD M3<T> add(const M3<T>& a, const M3<T>& b) {
  M3<T> c;
#pragma unroll
  for (int k = 0; k < 9; ++k) c.e[k] = a.e[k] + b.e[k];
  return c;
}
template <class T>
MAG_HD M3<T> sub(const M3<T>& a, const M3<T>& b) {
  M3<T> c;
#pragma unroll
  for (int k = 0; k < 9; ++k) c.e[k] = a.e[k] - b.e[k];
  return c;
}
template <class T>
MAG_HD void add_to(M3<T>& c, const M3<T>& a) {
#pragma unroll
  for (int k = 0; k < 9; ++k) c.e[k] = c.e[k] + a.e[k];
}
template <class T>
MAG_HD void add_dag_to(M3<T>& c, const M3<T>& a) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) c(i, j) = c(i, j) + cconj(a(j, i));
}
template <class T>
MAG_HD M3<T> scale(const M3<T>& a, T s) {
  M3<T> c;
#pragma unroll
  for (int k = 0; k < 9; ++k) c.e[k] = s * a.e[k];
  return c;
}
template <class T>
MAG_HD M3<T> cscale(const M3<T>& a, cpx<T> s) {
  M3<T> c;
#pragma unroll
  for (int k = 0; k < 9; ++k) c.e[k] = a.e[k] * s;
  return c;
}
// c += s * a
template <class T>
MAG_HD void axpy(M3<T>& c, T s, const M3<T>& a) {
#pragma unroll
  for (int k = 0; k < 9; ++k) c.e[k] = {mfma(s, a.e[k].r, c.e[k].r), mfma(s, a.e[k].i, c.e[k].i)};
}
template <class T>
MAG_HD void caxpy(M3<T>& c, cpx<T> s, const M3<T>& a) {
#pragma unroll
  for (int k = 0; k < 9; ++k) c.e[k] = cfma(s, a.e[k], c.e[k]);
}
template <class T>
MAG_HD cpx<T> trace(const M3<T>& a) { return a.e[0] + a.e[4] + a.e[8]; }
template <class T>
MAG_HD T retrace(const M3<T>& a) { return a.e[0].r + a.e[4].r + a.e[8].r; }
// Tr(A B)
template <class T>
MAG_HD cpx<T> trace_mul(const M3<T>& a, const M3<T>& b) {
  cpx<T> s = {T(0), T(0)};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) s = cfma(a(i, k), b(k, i), s);
  return s;
}
// Re Tr(A B)
template <class T>
MAG_HD T retrace_mul(const M3<T>& a, const M3<T>& b) {
  T s = T(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) s = mfma(a(i, k).r, b(k, i).r, mfma(-a(i, k).i, b(k, i).i, s));
  return s;
}
// Re Tr(A B^+) = sum_ij Re(A_ij conj(B_ij))
template <class T>
MAG_HD T retrace_mul_nd(const M3<T>& a, const M3<T>& b) {
  T s = T(0);
#pragma unroll
  for (int k = 0; k < 9; ++k) s = mfma(a.e[k].r, b.e[k].r, mfma(a.e[k].i, b.e[k].i, s));
  return s;
}

// -------------------------------------------------------------- su(3) algebra
// Hermitian traceless H = sum_a h_a T_a <-> 8 reals h_a = Tr(T_a H).
template <class T>
MAG_HD M3<T> alg_to_mat(const T h[8]) {
  const T s2 = T(0.70710678118654752440), s6 = T(0.40824829046386301637);
  M3<T> m;
  m(0, 1) = {s2 * h[0], -s2 * h[1]};
  m(1, 0) = {s2 * h[0], s2 * h[1]};
  m(0, 2) = {s2 * h[3], -s2 * h[4]};
  m(2, 0) = {s2 * h[3], s2 * h[4]};
  m(1, 2) = {s2 * h[5], -s2 * h[6]};
  m(2, 1) = {s2 * h[5], s2 * h[6]};
  m(0, 0) = {s2 * h[2] + s6 * h[7], T(0)};
  m(1, 1) = {-s2 * h[2] + s6 * h[7], T(0)};
  m(2, 2) = {T(-2) * s6 * h[7], T(0)};
  return m;
}
// h_a = Re Tr(T_a H) for any H (the anti-Hermitian and trace parts drop out).
template <class T>
MAG_HD void mat_to_alg(const M3<T>& m, T h[8]) {
  const T s2 = T(0.70710678118654752440), s6 = T(0.40824829046386301637);
  h[0] = s2 * (m(0, 1).r + m(1, 0).r);
  h[1] = s2 * (m(1, 0).i - m(0, 1).i);
  h[2] = s2 * (m(0, 0).r - m(1, 1).r);
  h[3] = s2 * (m(0, 2).r + m(2, 0).r);
  h[4] = s2 * (m(2, 0).i - m(0, 2).i);
  h[5] = s2 * (m(1, 2).r + m(2, 1).r);
  h[6] = s2 * (m(2, 1).i - m(1, 2).i);
  h[7] = s6 * (m(0, 0).r + m(1, 1).r - T(2) * m(2, 2).r);
}
// h_a of F = -i TA(W): F_a = Im Tr(T_a W).
template <class T>
MAG_HD void force_alg_from(const M3<T>& w, T h[8]) {
  const T s2 = T(0.70710678118654752440), s6 = T(0.40824829046386301637);
  h[0] = s2 * (w(0, 1).i + w(1, 0).i);
  h[1] = s2 * (w(0, 1).r - w(1, 0).r);
  h[2] = s2 * (w(0, 0).i - w(1, 1).i);
  h[3] = s2 * (w(0, 2).i + w(2, 0).i);
  h[4] = s2 * (w(0, 2).r - w(2, 0).r);
  h[5] = s2 * (w(1, 2).i + w(2, 1).i);
  h[6] = s2 * (w(1, 2).r - w(2, 1).r);
  h[7] = s6 * (w(0, 0).i + w(1, 1).i - T(2) * w(2, 2).i);
}
// TA(M) = (M - M^+)/2 - Tr(M - M^+)/6
template <class T>
MAG_HD M3<T> trless_antiherm(const M3<T>& m) {
  M3<T> a;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) a(i, j) = T(0.5) * (m(i, j) - cconj(m(j, i)));
  cpx<T> t = trace(a);
  t = {t.r / T(3), t.i / T(3)};
#pragma unroll
  for (int i = 0; i < 3; ++i) a(i, i) = a(i, i) - t;
  return a;
}
// HT(M) as 8 algebra components: (M + M^+)/2 projected = Re Tr(T_a M).
template <class T>
MAG_HD void herm_traceless_alg(const M3<T>& m, T h[8]) { mat_to_alg(m, h); }

// Two-row products. A product of SU(3) links is in SU(3), so only its rows 0
// and 1 need forming (two thirds of the multiply-adds); reconstruct_row2
// completes it when the third row is used. Row 2 of the result is zero.
// The first factor of mul_r2 / mul_nd_r2 is read only in rows 0 and 1.
// C = A B, rows 0-1
template <class T>
MAG_HD M3<T> mul_r2(const M3<T>& a, const M3<T>& b) {
  M3<T> c;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = a(i, 0) * b(0, j);
      s = cfma(a(i, 1), b(1, j), s);
      s = cfma(a(i, 2), b(2, j), s);
      c(i, j) = s;
    }
#pragma unroll
  for (int j = 0; j < 3; ++j) c(2, j) = {T(0), T(0)};
  return c;
}
// C = A B^+, rows 0-1
template <class T>
MAG_HD M3<T> mul_nd_r2(const M3<T>& a, const M3<T>& b) {
  M3<T> c;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = a(i, 0) * cconj(b(j, 0));
      s = cfma(a(i, 1), cconj(b(j, 1)), s);
      s = cfma(a(i, 2), cconj(b(j, 2)), s);
      c(i, j) = s;
    }
#pragma unroll
  for (int j = 0; j < 3; ++j) c(2, j) = {T(0), T(0)};
  return c;
}
// C = A^+ B, rows 0-1
template <class T>
MAG_HD M3<T> mul_dn_r2(const M3<T>& a, const M3<T>& b) {
  M3<T> c;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = cconj(a(0, i)) * b(0, j);
      s = cfma_cj(a(1, i), b(1, j), s);
      s = cfma_cj(a(2, i), b(2, j), s);
      c(i, j) = s;
    }
#pragma unroll
  for (int j = 0; j < 3; ++j) c(2, j) = {T(0), T(0)};
  return c;
}
// C = A^+ B^+, rows 0-1
template <class T>
MAG_HD M3<T> mul_dd_r2(const M3<T>& a, const M3<T>& b) {
  M3<T> c;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = cconj(a(0, i)) * cconj(b(j, 0));
      s = s + cconj(a(1, i)) * cconj(b(j, 1));
      s = s + cconj(a(2, i)) * cconj(b(j, 2));
      c(i, j) = s;
    }
#pragma unroll
  for (int j = 0; j < 3; ++j) c(2, j) = {T(0), T(0)};
  return c;
}

// Third row of an SU(3) matrix from the first two: conj(row0 x row1).
template <class T>
MAG_HD void reconstruct_row2(M3<T>& m) {
  m(2, 0) = cconj(m(0, 1) * m(1, 2) - m(0, 2) * m(1, 1));
  m(2, 1) = cconj(m(0, 2) * m(1, 0) - m(0, 0) * m(1, 2));
  m(2, 2) = cconj(m(0, 0) * m(1, 1) - m(0, 1) * m(1, 0));
}

// c += m for an SU(3) matrix m given by its rows 0 and 1
template <class T>
MAG_HD void add_su3_r2(M3<T>& c, M3<T> m) {
  reconstruct_row2(m);
#pragma unroll
  for (int k = 0; k < 9; ++k) c.e[k] = c.e[k] + m.e[k];
}
// c += s m for an SU(3) matrix m given by its rows 0 and 1
template <class T>
MAG_HD void axpy_su3_r2(M3<T>& c, T s, M3<T> m) {
  reconstruct_row2(m);
#pragma unroll
  for (int k = 0; k < 9; ++k) c.e[k] = {mfma(s, m.e[k].r, c.e[k].r), mfma(s, m.e[k].i, c.e[k].i)};
}

// Project onto SU(3): normalise row 0, orthogonalise and normalise row 1,
// row 2 = conj(row0 x row1) (the cps::Matrix::Unitarize rule).
template <class T>
MAG_HD void unitarize(M3<T>& m) {
  T n0 = T(0);
#pragma unroll
  for (int j = 0; j < 3; ++j) n0 += m(0, j).r * m(0, j).r + m(0, j).i * m(0, j).i;
  n0 = T(1) / sqrt(n0);
#pragma unroll
  for (int j = 0; j < 3; ++j) m(0, j) = n0 * m(0, j);
  cpx<T> c = {T(0), T(0)};
#pragma unroll
  for (int j = 0; j < 3; ++j) c = c + cconj(m(0, j)) * m(1, j);
#pragma unroll
  for (int j = 0; j < 3; ++j) m(1, j) = m(1, j) - c * m(0, j);
  T n1 = T(0);
#pragma unroll
  for (int j = 0; j < 3; ++j) n1 += m(1, j).r * m(1, j).r + m(1, j).i * m(1, j).i;
  n1 = T(1) / sqrt(n1);
#pragma unroll
  for (int j = 0; j < 3; ++j) m(1, j) = n1 * m(1, j);
  reconstruct_row2(m);
}

}  // namespace mag
