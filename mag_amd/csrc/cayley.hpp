// Cayley-Hamilton exponential exp(iQ) of a Hermitian traceless 3x3 matrix and
// its derivative coefficients (Morningstar & Peardon, hep-lat/0311018,
// eqs. 17-34 and 69-70), as the reference codes them in expiQ
// (field-hmc.h:82-123), DexpiQ (field-multi.h:90-178) and compute_Lambda
// (field-multi.h:180-269).
//
// exp(iQ) = f0 + f1 Q + f2 Q^2. The scalar coefficients are always evaluated
// in FP64 from the invariants c0 = Re Tr Q^3 / 3 and c1 = Re Tr Q^2 / 2, also
// on the FP32 molecular-dynamics path: the h_j and b_ij formulas cancel O(1)
// terms down to O(u^2) (and b_ij divide by (9u^2 - w^2)^2), which FP32 cannot
// carry at the small |Q| of force-gradient sub-steps. The 3x3 algebra around
// them stays in the kernel's own precision.
//
// The acos argument c0/c0max is clamped to 1: rounding can push it past 1,
// where the reference formula returns NaN.
#pragma once
#include "su3.hpp"

namespace mag {

struct CHCoef {
  bool zero;
  cpx<double> f0, f1, f2;
  cpx<double> b10, b11, b12, b20, b21, b22;
};

MAG_HD cpx<double> cexpi(double x) { return {cos(x), sin(x)}; }

template <bool NEED_B>
MAG_HD CHCoef ch_coefficients(double c0, double c1) {
  CHCoef c;
  bool reflect = false;
  if (c0 < 0.) { c0 = -c0; reflect = true; }
  c.zero = (c1 == 0.);
  if (c.zero) {
    c.f0 = {1., 0.}; c.f1 = {0., 0.}; c.f2 = {0., 0.};
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
  return c;
}

// Invariants of a Hermitian traceless Q, accumulated in FP64.
template <class T>
MAG_HD void ch_invariants(const M3<T>& Q, const M3<T>& Q2, double& c0, double& c1) {
  double t2 = 0., t3 = 0.;
#pragma unroll
  for (int k = 0; k < 3; ++k) t2 += double(Q2.e[4 * k].r);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      t3 += double(Q2(i, k).r) * double(Q(k, i).r) - double(Q2(i, k).i) * double(Q(k, i).i);
  c0 = t3 / 3.;
  c1 = t2 / 2.;
}

template <class T>
MAG_HD cpx<T> to(cpx<double> a) { return {T(a.r), T(a.i)}; }

// f0 + f1 Q + f2 Q^2
template <class T>
MAG_HD M3<T> ch_combine(const M3<T>& Q, const M3<T>& Q2, cpx<double> f0, cpx<double> f1, cpx<double> f2) {
  const cpx<T> g0 = to<T>(f0), g1 = to<T>(f1), g2 = to<T>(f2);
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
  const M3<T> Q2 = mul(Q, Q);
  double c0, c1;
  ch_invariants(Q, Q2, c0, c1);
  CHCoef c = ch_coefficients<false>(c0, c1);
  if (c.zero) return unit3<T>();
  return ch_combine(Q, Q2, c.f0, c.f1, c.f2);
}

// Lambda of compute_Lambda (field-multi.h:180-269), returned as its 8
// algebra components, given Q and X = U SigmaP (so Tr(SigmaP B U) = Tr(X B)).
// Also returns exp(iQ) for the chain's Z' e^{iQ} term.
template <class T>
MAG_HD void ch_lambda(const M3<T>& Q, const M3<T>& X, T lam[8], M3<T>& eQ) {
  const M3<T> Q2 = mul(Q, Q);
  double c0, c1;
  ch_invariants(Q, Q2, c0, c1);
  CHCoef c = ch_coefficients<true>(c0, c1);
  if (c.zero) {
#pragma unroll
    for (int a = 0; a < 8; ++a) lam[a] = T(0);
    eQ = unit3<T>();
    return;
  }
  eQ = ch_combine(Q, Q2, c.f0, c.f1, c.f2);
  const M3<T> B1 = ch_combine(Q, Q2, c.b10, c.b11, c.b12);
  const M3<T> B2 = ch_combine(Q, Q2, c.b20, c.b21, c.b22);
  const cpx<T> t1 = trace_mul(X, B1), t2 = trace_mul(X, B2);
  // Gamma = t1 Q + t2 Q^2 + f1 X + f2 (Q X + X Q)
  const cpx<T> g1 = to<T>(c.f1), g2 = to<T>(c.f2);
  M3<T> QX = mul(Q, X);
  mul_acc(QX, X, Q);
  M3<T> G;
#pragma unroll
  for (int k = 0; k < 9; ++k) G.e[k] = cfma(t1, Q.e[k], cfma(t2, Q2.e[k], cfma(g1, X.e[k], g2 * QX.e[k])));
  mat_to_alg(G, lam);  // HT(Gamma) in generator components
}

}  // namespace mag
