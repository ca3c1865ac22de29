// Cayley-Hamilton exponential exp(iQ) of a Hermitian traceless 3x3 matrix and
// its derivative coefficients (Morningstar & Peardon, hep-lat/0311018,
// eqs. 17-34 and 69-70), as the reference codes them in expiQ
// (field-hmc.h:82-123), DexpiQ (field-multi.h:90-178) and compute_Lambda
// (field-multi.h:180-269).
//
// exp(iQ) = f0 + f1 Q + f2 Q^2. On the FP64 path the scalar coefficients
// are the closed forms of the reference, evaluated from the invariants
// c0 = Re Tr Q^3 / 3 and c1 = Re Tr Q^2 / 2. The acos argument c0/c0max is
// clamped to 1: rounding can push it past 1, where the reference formula
// returns NaN.
//
// The FP32 molecular-dynamics path cannot use those closed forms: h_j cancels
// O(1) terms down to O(u^2) and b_ij divide by (9u^2 - w^2)^2 ~ c1^2, so at
// the small |Q| of force-gradient sub-steps FP32 loses every digit (FP64
// itself keeps only ~5 digits of b_2j at c1 = 1e-6). There the coefficients
// come from the exponential series reduced with Cayley-Hamilton
// (Q^3 = c1 Q + c0 for traceless Hermitian Q): with
// (iQ)^n / n! = i^n (a_n + b_n Q + d_n Q^2),
//   a_{n+1} = c0 d_n / (n+1),  b_{n+1} = (a_n + c1 d_n) / (n+1),
//   d_{n+1} = b_n / (n+1),     (a_0, b_0, d_0) = (1, 0, 0),
// all real, f_j = sum_n i^n (a_n, b_n, d_n)_j, and b_1j = df_j/dc1,
// b_2j = df_j/dc0 (hep-lat/0311018 eq. 57) from the recurrence differentiated
// term by term. 14 terms hold the truncation below 1e-10 for c1 <= 1 (largest
// eigenvalue |q| <= sqrt(4 c1 / 3)); larger Q fall back to the FP64 closed
// forms. Both branches keep the reference's exact-zero case (c1 == 0 gives
// exp(iQ) = 1 and Lambda = 0).
#pragma once
#include "su3.hpp"

namespace mag {

template <class T>
struct CHC {
  bool zero;
  cpx<T> f0, f1, f2;
  cpx<T> f0m1;  // f0 - 1, carried separately for exp(iQ) - 1
  cpx<T> b10, b11, b12, b20, b21, b22;
};
using CHCoef = CHC<double>;

MAG_HD cpx<double> cexpi(double x) { return {cos(x), sin(x)}; }

template <bool NEED_B>
MAG_HD CHCoef ch_coefficients(double c0, double c1) {
  CHCoef c;
  bool reflect = false;
  if (c0 < 0.) { c0 = -c0; reflect = true; }
  c.zero = (c1 == 0.);
  if (c.zero) {
    c.f0 = {1., 0.}; c.f1 = {0., 0.}; c.f2 = {0., 0.}; c.f0m1 = {0., 0.};
    c.b10 = c.b11 = c.b12 = c.b20 = c.b21 = c.b22 = {0., 0.};
    return c;
  }
  const double c13 = c1 / 3.;
  const double c0m = 2. * c13 * sqrt(c13);  // 2 (c1/3)^{3/2}
  double ratio = c0 / c0m;
  if (ratio > 1.) ratio = 1.;
  const double theta = acos(ratio);
  const double u = sqrt(c13) * cos(theta / 3.);
  const double w = sqrt(c1) * sin(theta / 3.);
  const double uu = u * u, ww = w * w;
  double xi0, xi1 = 0.;
  const double sw = sin(w), cw = cos(w);
  if (ww < 0.05 * 0.05) {
    xi0 = 1. - ww / 6. * (1. - ww / 20. * (1. - ww / 42.));
    if (NEED_B) xi1 = -0.5 * (1. - ww / 12. * (1. - ww / 30.)) + 1. / 6. * (1. - ww / 20. * (1. - ww / 42.));
  } else {
    xi0 = sw / w;
    if (NEED_B) xi1 = cw / ww - sw / (ww * w);
  }
  const cpx<double> e2u = cexpi(2. * u), emu = cexpi(-u);
  const cpx<double> I = {0., 1.};
  // h0 = (u^2 - w^2) e^{2iu} + e^{-iu} (8 u^2 cos w + 2 i u (3u^2 + w^2) xi0)
  cpx<double> h0 = (uu - ww) * e2u + emu * cpx<double>{8. * uu * cw, 2. * u * (3. * uu + ww) * xi0};
  // h1 = 2u e^{2iu} - e^{-iu} (2u cos w - i (3u^2 - w^2) xi0)
  cpx<double> h1 = (2. * u) * e2u - emu * cpx<double>{2. * u * cw, -(3. * uu - ww) * xi0};
  // h2 = e^{2iu} - e^{-iu} (cos w + 3 i u xi0)
  cpx<double> h2 = e2u - emu * cpx<double>{cw, 3. * u * xi0};
  const double den = 9. * uu - ww;
  const double rden = 1. / den;
  c.f0 = rden * h0; c.f1 = rden * h1; c.f2 = rden * h2;
  if (NEED_B) {
    // r_ij of hep-lat/0311018 eqs. 60-65 in the reference's grouping.
    cpx<double> r01 = 2. * (cpx<double>{u, uu - ww} * e2u) +
                      2. * (emu * (4. * u * cw * cpx<double>{2., -u} +
                                   I * (xi0 * cpx<double>{9. * uu + ww, -u * (3. * uu + ww)})));
    cpx<double> r11 = 2. * (cpx<double>{1., 2. * u} * e2u) +
                      emu * ((-2. * cw) * cpx<double>{1., -u} + I * (xi0 * cpx<double>{6. * u, ww - 3. * uu}));
    cpx<double> r21 = 2. * (I * e2u) + I * (emu * (cpx<double>{cw, 0.} - 3. * xi0 * cpx<double>{1., -u}));
    cpx<double> r02 = -2. * e2u + (2. * u) * (I * (emu * (cpx<double>{cw, 0.} + xi0 * cpx<double>{1., 4. * u} +
                                                          cpx<double>{3. * uu * xi1, 0.})));
    cpx<double> r12 = cpx<double>{0., -1.} *
                      (emu * (cpx<double>{cw, 0.} + xi0 * cpx<double>{1., 2. * u} - cpx<double>{3. * uu * xi1, 0.}));
    cpx<double> r22 = emu * cpx<double>{xi0, -3. * u * xi1};
    const double d2 = 1. / (2. * den * den);
    const double a1 = 3. * uu - ww, a2 = 2. * (15. * uu + ww);
    c.b10 = d2 * ((2. * u) * r01 + a1 * r02 - a2 * c.f0);
    c.b11 = d2 * ((2. * u) * r11 + a1 * r12 - a2 * c.f1);
    c.b12 = d2 * ((2. * u) * r21 + a1 * r22 - a2 * c.f2);
    c.b20 = d2 * (r01 - (3. * u) * r02 - (24. * u) * c.f0);
    c.b21 = d2 * (r11 - (3. * u) * r12 - (24. * u) * c.f1);
    c.b22 = d2 * (r21 - (3. * u) * r22 - (24. * u) * c.f2);
  }
  if (reflect) {
    c.f0 = cconj(c.f0);
    c.f1 = {-c.f1.r, c.f1.i};
    c.f2 = cconj(c.f2);
    if (NEED_B) {
      c.b10 = cconj(c.b10);
      c.b11 = {-c.b11.r, c.b11.i};
      c.b12 = cconj(c.b12);
      c.b20 = {-c.b20.r, c.b20.i};
      c.b21 = cconj(c.b21);
      c.b22 = {-c.b22.r, c.b22.i};
    }
  }
  c.f0m1 = {c.f0.r - 1., c.f0.i};
  return c;
}

constexpr int CH_SERIES_TERMS = 14;
constexpr float CH_SERIES_C1_MAX = 1.0f;

// i^n x added to (re, im), n known at compile time after unrolling.
MAG_HD void add_phase(int n, float x, float& re, float& im) {
  switch (n & 3) {
    case 0: re += x; break;
    case 1: im += x; break;
    case 2: re -= x; break;
    default: im -= x; break;
  }
}

// Series coefficients for c1 <= CH_SERIES_C1_MAX (FP32 path).
template <bool NEED_B>
MAG_HD CHC<float> ch_series(float c0, float c1) {
  CHC<float> c;
  c.zero = (c1 == 0.f);
  float fr[3] = {0.f, 0.f, 0.f}, fi[3] = {0.f, 0.f, 0.f};  // fr[0]: f0 - 1
  float pr[3] = {0.f, 0.f, 0.f}, pi[3] = {0.f, 0.f, 0.f};
  float qr[3] = {0.f, 0.f, 0.f}, qi[3] = {0.f, 0.f, 0.f};
  float a = 1.f, b = 0.f, d = 0.f;                  // (iQ)^n / n! in units of i^n
  float pa = 0.f, pb = 0.f, pd = 0.f;               // d/dc1
  float qa = 0.f, qb = 0.f, qd = 0.f;               // d/dc0
#pragma unroll
  for (int n = 1; n < CH_SERIES_TERMS; ++n) {
    const float inv = 1.f / float(n);
    if (NEED_B) {
      const float npa = c0 * pd * inv, npb = (pa + c1 * pd + d) * inv, npd = pb * inv;
      const float nqa = (c0 * qd + d) * inv, nqb = (qa + c1 * qd) * inv, nqd = qb * inv;
      pa = npa; pb = npb; pd = npd;
      qa = nqa; qb = nqb; qd = nqd;
      add_phase(n, pa, pr[0], pi[0]); add_phase(n, pb, pr[1], pi[1]); add_phase(n, pd, pr[2], pi[2]);
      add_phase(n, qa, qr[0], qi[0]); add_phase(n, qb, qr[1], qi[1]); add_phase(n, qd, qr[2], qi[2]);
    }
    const float na = c0 * d * inv, nb = (a + c1 * d) * inv, nd = b * inv;
    a = na; b = nb; d = nd;
    add_phase(n, a, fr[0], fi[0]); add_phase(n, b, fr[1], fi[1]); add_phase(n, d, fr[2], fi[2]);
  }
  if (c.zero) {
    c.f0 = {1.f, 0.f}; c.f1 = {0.f, 0.f}; c.f2 = {0.f, 0.f}; c.f0m1 = {0.f, 0.f};
    c.b10 = c.b11 = c.b12 = c.b20 = c.b21 = c.b22 = {0.f, 0.f};
    return c;
  }
  c.f0m1 = {fr[0], fi[0]};
  c.f0 = {1.f + fr[0], fi[0]}; c.f1 = {fr[1], fi[1]}; c.f2 = {fr[2], fi[2]};
  c.b10 = {pr[0], pi[0]}; c.b11 = {pr[1], pi[1]}; c.b12 = {pr[2], pi[2]};
  c.b20 = {qr[0], qi[0]}; c.b21 = {qr[1], qi[1]}; c.b22 = {qr[2], qi[2]};
  return c;
}

template <class T>
MAG_HD CHC<T> to_coef(const CHC<double>& c) {
  CHC<T> r;
  r.zero = c.zero;
  r.f0 = {T(c.f0.r), T(c.f0.i)}; r.f1 = {T(c.f1.r), T(c.f1.i)}; r.f2 = {T(c.f2.r), T(c.f2.i)};
  r.f0m1 = {T(c.f0m1.r), T(c.f0m1.i)};
  r.b10 = {T(c.b10.r), T(c.b10.i)}; r.b11 = {T(c.b11.r), T(c.b11.i)}; r.b12 = {T(c.b12.r), T(c.b12.i)};
  r.b20 = {T(c.b20.r), T(c.b20.i)}; r.b21 = {T(c.b21.r), T(c.b21.i)}; r.b22 = {T(c.b22.r), T(c.b22.i)};
  return r;
}

// Invariants of a Hermitian traceless Q in the kernel's precision.
template <class T>
MAG_HD void ch_invariants(const M3<T>& Q, const M3<T>& Q2, T& c0, T& c1) {
  T t2 = T(0), t3 = T(0);
#pragma unroll
  for (int k = 0; k < 3; ++k) t2 += Q2.e[4 * k].r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) t3 += Q2(i, k).r * Q(k, i).r - Q2(i, k).i * Q(k, i).i;
  c0 = t3 / T(3);
  c1 = t2 / T(2);
}

// Coefficients for the kernel's precision: FP64 closed forms, or on the FP32
// path the series where it converges.
template <bool NEED_B, class T>
MAG_HD CHC<T> ch_coef(T c0, T c1) {
  if constexpr (sizeof(T) == sizeof(float)) {
    if (c1 <= CH_SERIES_C1_MAX) return ch_series<NEED_B>(c0, c1);
  }
  return to_coef<T>(ch_coefficients<NEED_B>(double(c0), double(c1)));
}

template <class T>
MAG_HD cpx<T> to(cpx<double> a) { return {T(a.r), T(a.i)}; }

// f0 + f1 Q + f2 Q^2
template <class T>
MAG_HD M3<T> ch_combine(const M3<T>& Q, const M3<T>& Q2, cpx<T> g0, cpx<T> g1, cpx<T> g2) {
  M3<T> r;
#pragma unroll
  for (int k = 0; k < 9; ++k) r.e[k] = cfma(g2, Q2.e[k], g1 * Q.e[k]);
#pragma unroll
  for (int i = 0; i < 3; ++i) r(i, i) = r(i, i) + g0;
  return r;
}

// exp(iQ) (field-hmc.h:82-123).
template <class T>
MAG_HD M3<T> expiQ(const M3<T>& Q) {
  const M3<T> Q2 = sq_h(Q);
  T c0, c1;
  ch_invariants(Q, Q2, c0, c1);
  const CHC<T> c = ch_coef<false>(c0, c1);
  if (c.zero) return unit3<T>();
  return ch_combine(Q, Q2, c.f0, c.f1, c.f2);
}

// exp(iQ) - 1 = (f0 - 1) + f1 Q + f2 Q^2, never forming 1 + (small).
template <class T>
MAG_HD M3<T> expiQ_m1(const M3<T>& Q) {
  const M3<T> Q2 = sq_h(Q);
  T c0, c1;
  ch_invariants(Q, Q2, c0, c1);
  const CHC<T> c = ch_coef<false>(c0, c1);
  M3<T> r;
#pragma unroll
  for (int k = 0; k < 9; ++k) r.e[k] = cfma(c.f2, Q2.e[k], c.f1 * Q.e[k]);
  if (c.zero) return zero3<T>();
#pragma unroll
  for (int i = 0; i < 3; ++i) r(i, i) = r(i, i) + c.f0m1;
  return r;
}

// Link updates exp(iQ) U and U exp(iQ). FP64 multiplies by exp(iQ), as the
// reference does. FP32 adds (exp(iQ) - 1) U to U: forming exp(iQ) first
// rounds the O(|Q|^2) parts of f0 = 1 - c1/2 + ... against 1, and over a
// trajectory that shrinks every link by ~2e-7 in the same direction (a
// systematic shift of the FP32 dH); U + (exp(iQ) - 1) U rounds once per
// element with no bias.
template <class T>
MAG_HD M3<T> exp_left(const M3<T>& Q, const M3<T>& U) {
  if constexpr (sizeof(T) == sizeof(float)) {
    M3<T> r = U;
    mul_acc(r, expiQ_m1(Q), U);
    return r;
  } else {
    return mul(expiQ(Q), U);
  }
}
template <class T>
MAG_HD M3<T> exp_right(const M3<T>& U, const M3<T>& Q) {
  if constexpr (sizeof(T) == sizeof(float)) {
    M3<T> r = U;
    mul_acc(r, U, expiQ_m1(Q));
    return r;
  } else {
    return mul(U, expiQ(Q));
  }
}

// DexpiQ (field-multi.h:90-178): dE = d exp(iQ(t))/dt for dQ/dt = DQ, from
// B1, B2 (MP eqs. 69-70): dE = Tr(Q DQ) B1 + Tr(Q^2 DQ) B2 + f1 DQ
// + f2 (DQ Q + Q DQ), the B terms folded into coefficients of 1, Q, Q^2.
// Q is Hermitian traceless. Zero when c1 == 0 (exactly Q = 0), as in the
// reference. E = exp(iQ) comes with it.
template <class T>
MAG_HD void dexpiq(const M3<T>& Q, const M3<T>& DQ, M3<T>& E, M3<T>& dE) {
  const M3<T> Q2 = mul(Q, Q);
  T c0, c1;
  ch_invariants(Q, Q2, c0, c1);
  const CHC<T> c = ch_coef<true>(c0, c1);
  if (c.zero) {
    E = unit3<T>();
    dE = zero3<T>();
    return;
  }
  E = ch_combine(Q, Q2, c.f0, c.f1, c.f2);
  const cpx<T> t1 = trace_mul(Q, DQ), t2 = trace_mul(Q2, DQ);
  dE = ch_combine(Q, Q2, t1 * c.b10 + t2 * c.b20, t1 * c.b11 + t2 * c.b21, t1 * c.b12 + t2 * c.b22);
  M3<T> QD = mul_hn(Q, DQ);
  mul_acc_nh(QD, DQ, Q);  // DQ Q + Q DQ
#pragma unroll
  for (int k = 0; k < 9; ++k) dE.e[k] = cfma(c.f1, DQ.e[k], cfma(c.f2, QD.e[k], dE.e[k]));
}

// Lambda of compute_Lambda (field-multi.h:180-269), returned as its 8
// algebra components, given Q and X = U SigmaP (so Tr(SigmaP B U) = Tr(X B)).
// Also returns exp(iQ) for the chain's Z' e^{iQ} term.
template <class T>
MAG_HD void ch_lambda(const M3<T>& Q, const M3<T>& X, T lam[8], M3<T>& eQ) {
  const M3<T> Q2 = sq_h(Q);
  T c0, c1;
  ch_invariants(Q, Q2, c0, c1);
  const CHC<T> c = ch_coef<true>(c0, c1);
  if (c.zero) {
#pragma unroll
    for (int a = 0; a < 8; ++a) lam[a] = T(0);
    eQ = unit3<T>();
    return;
  }
  eQ = ch_combine(Q, Q2, c.f0, c.f1, c.f2);
  // Tr(X B_j) with B_j = b_j0 + b_j1 Q + b_j2 Q^2, from Tr X, Tr(X Q), Tr(X Q^2)
  const cpx<T> x0 = trace(X), x1 = trace_mul_nh(X, Q), x2 = trace_mul_nh(X, Q2);
  const cpx<T> t1 = cfma(c.b12, x2, cfma(c.b11, x1, c.b10 * x0));
  const cpx<T> t2 = cfma(c.b22, x2, cfma(c.b21, x1, c.b20 * x0));
  // Gamma = t1 Q + t2 Q^2 + f1 X + f2 (Q X + X Q)
  const cpx<T> g1 = c.f1, g2 = c.f2;
  M3<T> QX = mul_hn(Q, X);
  mul_acc_nh(QX, X, Q);
  M3<T> G;
#pragma unroll
  for (int k = 0; k < 9; ++k) G.e[k] = cfma(t1, Q.e[k], cfma(t2, Q2.e[k], cfma(g1, X.e[k], g2 * QX.e[k])));
  mat_to_alg(G, lam);  // HT(Gamma) in generator components
}

}  // namespace mag
