// Host-side check of the FP32 Cayley-Hamilton series coefficients
// (mag_amd/csrc/cayley.hpp, ch_series) against the FP64 closed forms of
// hep-lat/0311018 (ch_coefficients) on random traceless Hermitian spectra
// with c1 from 1e-3 to 1, where the closed forms keep >= 9 digits. Runs on
// the CPU: the device headers are host-callable. Prints the worst absolute
// error of f_j, b_1j, b_2j and exits non-zero above the bound.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../mag_amd/csrc/cayley.hpp"

using namespace mag;

static double err(cpx<float> a, cpx<double> b) { return std::hypot(double(a.r) - b.r, double(a.i) - b.i); }

int main() {
  std::mt19937_64 gen(7);
  std::normal_distribution<double> nd;
  double ef = 0., e1 = 0., e2 = 0.;
  for (double c1 : {1e-3, 1e-2, 0.1, 0.5, 1.0}) {
    for (int t = 0; t < 200; ++t) {
      double q[3] = {nd(gen), nd(gen), nd(gen)};
      const double m = (q[0] + q[1] + q[2]) / 3.;
      for (double& x : q) x -= m;
      const double s = std::sqrt(2. * c1 / (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]));
      for (double& x : q) x *= s;
      const double c0 = q[0] * q[1] * q[2];  // det Q = Tr Q^3 / 3
      const CHC<double> a = ch_coefficients<true>(c0, c1);
      const CHC<float> b = ch_series<true>(float(c0), float(c1));
      ef = std::fmax(ef, std::fmax(err(b.f0, a.f0), std::fmax(err(b.f1, a.f1), err(b.f2, a.f2))));
      e1 = std::fmax(e1, std::fmax(err(b.b10, a.b10), std::fmax(err(b.b11, a.b11), err(b.b12, a.b12))));
      e2 = std::fmax(e2, std::fmax(err(b.b20, a.b20), std::fmax(err(b.b21, a.b21), err(b.b22, a.b22))));
    }
  }
  // exp(iQ) at a known spectrum: Q = diag(q) gives f0 + f1 q + f2 q^2 = e^{iq}.
  const double q[3] = {0.4, -0.1, -0.3};
  const double c1 = (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) / 2., c0 = q[0] * q[1] * q[2];
  const CHC<float> b = ch_series<false>(float(c0), float(c1));
  double ed = 0.;
  for (double x : q) {
    const double re = b.f0.r + b.f1.r * x + b.f2.r * x * x, im = b.f0.i + b.f1.i * x + b.f2.i * x * x;
    ed = std::fmax(ed, std::hypot(re - std::cos(x), im - std::sin(x)));
  }
  // exp(iQ) - 1 at small |Q| (the FP32 link update U + (exp(iQ) - 1) U):
  // relative to |e^{iq} - 1| on a diagonal Q, e^{iq} - 1 = (-2 sin^2(q/2), sin q).
  double em = 0.;
  for (double scale : {1e-5, 1e-3, 1e-1}) {
    const double d[3] = {0.7 * scale, -0.2 * scale, -0.5 * scale};
    const double k1 = (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) / 2., k0 = d[0] * d[1] * d[2];
    const CHC<float> m = ch_series<false>(float(k0), float(k1));
    for (double x : d) {
      const double re = m.f0m1.r + m.f1.r * x + m.f2.r * x * x, im = m.f0m1.i + m.f1.i * x + m.f2.i * x * x;
      const double sh = std::sin(x / 2.);
      em = std::fmax(em, std::hypot(re + 2. * sh * sh, im - std::sin(x)) / std::fabs(x));
    }
  }
  std::printf("f %.3e  b1 %.3e  b2 %.3e  diag %.3e  expm1 %.3e\n", ef, e1, e2, ed, em);
  return (ef < 2e-7 && e1 < 2e-7 && e2 < 2e-7 && ed < 2e-7 && em < 1e-6) ? 0 : 1;
}
