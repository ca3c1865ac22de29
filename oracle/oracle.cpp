// ORACLE — test infrastructure only. FP64 CPU restatement of the reference's
// constrained-HMC hot path (hummingtree/mag: field-hmc.h, field-matrix.h,
// field-multi.h, field-staple.h, field-demon.h, field-wloop.h). It is the
// checker that the GPU library is compared against; only tests/, the smoke
// check and the benchmark's CPU-baseline leg load it. The shipped library
// never links or calls it.
//
// Loop structure, path lists and operation order follow the reference; each
// function cites the reference lines it restates. The reference itself cannot
// be compiled (qlat, CPS, Timer are absent), so CPS/qlat semantics are
// restated from their definitions (see su3.hpp). Deliberate deviations:
//  * RNG: qlat's SHA-256 RngState is replaced by Philox4x32-10 streams keyed
//    by the chain seed (counter = {index, draw, epoch, stream}); the GPU
//    library uses the same streams.
//  * expiQ/DexpiQ/compute_Lambda clamp c0/c0max to [-1, 1] before acos; the
//    reference returns NaN when rounding pushes the ratio past 1.
//  * Reductions sum per-site values serially in site order (deterministic
//    for any thread count) instead of per-thread partial sums.
//  * A NaN Hamiltonian is rejected even during forced-accept trajectories.
//
// Every function is pinned by a check outside this file (tests/test_oracle.py,
// tests/test_oracle_pins.py, tests/test_multi_oracle.py,
// tests/test_local_oracle.py): scipy expm, Random123 known answers, finite
// differences, independent numpy restatements (lattice sums, the stout forward
// map and get_Q of hep-lat/0311018 eqs. 1-3, brute-force Wilson loops), the
// chair/twist tables' endpoint structure, invariants and closed forms. Parity
// unpinned: the qlat random streams this file replaces with Philox (no qlat
// here), and CPS operation order at the rounding level.
#include "oracle.h"
#include "su3.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace orc;

namespace {

constexpr int DIMN = 4;

// ---------------------------------------------------------------- geometry
struct Geo {
  int L[4];
  long V;
  explicit Geo(const int* l) {
    V = 1;
    for (int i = 0; i < 4; ++i) { L[i] = l[i]; V *= l[i]; }
  }
  long index(const int* x) const {
    int y[4];
    for (int i = 0; i < 4; ++i) y[i] = ((x[i] % L[i]) + L[i]) % L[i];
    return y[0] + (long)L[0] * (y[1] + (long)L[1] * (y[2] + (long)L[2] * y[3]));
  }
  void coord(long s, int* x) const {
    for (int i = 0; i < 4; ++i) { x[i] = (int)(s % L[i]); s /= L[i]; }
  }
};

struct Field {
  const Geo* g;
  Mat* m;  // [V][4]
  Mat& at(const int* x, int mu) const { return m[g->index(x) * 4 + mu]; }
  Mat& at(long s, int mu) const { return m[s * 4 + mu]; }
};

std::vector<Mat> load_field(const Geo& g, const double* p) {
  std::vector<Mat> f(g.V * 4);
  for (long k = 0; k < g.V * 4; ++k) f[k].load(p + 18 * k);
  return f;
}
void store_field(const std::vector<Mat>& f, double* p) {
  for (size_t k = 0; k < f.size(); ++k) f[k].store(p + 18 * k);
}

// ---------------------------------------------------------------- generators
// field-hmc.h:30-71: T_a = lambda_a / sqrt(2), Tr(T_a T_b) = delta_ab.
std::vector<Mat> make_generators() {
  const double is2 = 1. / std::sqrt(2.);
  std::vector<Mat> T(8);
  T[0](0, 1) = cplx(1.); T[0](1, 0) = cplx(1.);
  T[1](0, 1) = cplx(0., -1.); T[1](1, 0) = cplx(0., 1.);
  T[2](0, 0) = cplx(1.); T[2](1, 1) = cplx(-1.);
  T[3](0, 2) = cplx(1.); T[3](2, 0) = cplx(1.);
  T[4](0, 2) = cplx(0., -1.); T[4](2, 0) = cplx(0., 1.);
  T[5](1, 2) = cplx(1.); T[5](2, 1) = cplx(1.);
  T[6](1, 2) = cplx(0., -1.); T[6](2, 1) = cplx(0., 1.);
  for (int a = 0; a < 7; ++a) T[a] = T[a] * is2;
  T[7](0, 0) = cplx(1.); T[7](1, 1) = cplx(1.); T[7](2, 2) = cplx(-2.);
  T[7] = T[7] * (1. / std::sqrt(6.));
  return T;
}
const std::vector<Mat>& gens() {
  static const std::vector<Mat> T = make_generators();
  return T;
}

// ---------------------------------------------------------------- exp(iQ)
// Cayley-Hamilton coefficients of hep-lat/0311018 as coded in
// field-hmc.h:82-123 (f_j) and field-multi.h:90-178 (r_ij, b_ij).
struct CH {
  bool zero, reflect;
  double u, w, xi0, xi1;
  cplx f0, f1, f2, b10, b11, b12, b20, b21, b22;
};

CH ch_coeffs(const Mat& Q, bool need_b) {
  CH c{};
  double c0 = retr(Q * Q * Q) / 3.;
  c.reflect = false;
  if (c0 < 0.) { c0 = -c0; c.reflect = true; }
  double c1 = retr(Q * Q) / 2.;
  c.zero = (c1 == 0.);
  if (c.zero) return c;
  double c0m = 2. * std::pow(c1 / 3., 1.5);
  double ratio = c0 / c0m;
  if (ratio > 1.) ratio = 1.;  // deviation: clamp (the reference would return NaN)
  double theta = std::acos(ratio);
  double u = std::sqrt(c1 / 3.) * std::cos(theta / 3.);
  double w = std::sqrt(c1) * std::sin(theta / 3.);
  double xi0, xi1 = 0.;
  if (w * w < 0.05 * 0.05) {
    xi0 = 1. - w * w / 6. * (1. - w * w / 20. * (1. - w * w / 42.));
    if (need_b)
      xi1 = -1. / 2. * (1. - w * w / 12. * (1. - w * w / 30.)) +
            1. / 6. * (1. - w * w / 20. * (1. - w * w / 42.));
  } else {
    xi0 = std::sin(w) / w;
    if (need_b) xi1 = std::cos(w) / (w * w) - std::sin(w) / (w * w * w);
  }
  const cplx e2u = expix(2. * u), emu = expix(-u);
  const double cw = std::cos(w);
  cplx h0 = (u * u - w * w) * e2u + emu * (cplx(8. * u * u * cw) + I() * (2. * u * (3. * u * u + w * w) * xi0));
  cplx h1 = (2. * u) * e2u - emu * (cplx(2. * u * cw) - I() * ((3. * u * u - w * w) * xi0));
  cplx h2 = e2u - emu * (cplx(cw) + I() * (3. * u * xi0));
  const double den = 9. * u * u - w * w;
  c.f0 = h0 / den; c.f1 = h1 / den; c.f2 = h2 / den;
  c.u = u; c.w = w; c.xi0 = xi0; c.xi1 = xi1;
  if (need_b) {
    cplx r01 = 2. * (cplx(u) + I() * (u * u - w * w)) * e2u +
               2. * emu * (4. * u * (cplx(2.) - I() * u) * cw +
                           I() * xi0 * (cplx(9. * u * u + w * w) - I() * (u * (3. * u * u + w * w))));
    cplx r11 = 2. * (cplx(1.) + 2. * I() * u) * e2u +
               emu * (-2. * (cplx(1.) - I() * u) * cw + I() * xi0 * (cplx(6. * u) + I() * (w * w - 3. * u * u)));
    cplx r21 = 2. * I() * e2u + I() * emu * (cplx(cw) - 3. * xi0 * (cplx(1.) - I() * u));
    cplx r02 = -2. * e2u + 2. * I() * u * emu * (cplx(cw) + (cplx(1.) + 4. * I() * u) * xi0 + cplx(3. * u * u * xi1));
    cplx r12 = -I() * emu * (cplx(cw) + (cplx(1.) + 2. * I() * u) * xi0 - cplx(3. * u * u * xi1));
    cplx r22 = emu * (cplx(xi0) - 3. * I() * u * xi1);
    const double d2 = 2. * den * den;
    c.b10 = (2. * u * r01 + (3. * u * u - w * w) * r02 - 2. * (15. * u * u + w * w) * c.f0) / d2;
    c.b11 = (2. * u * r11 + (3. * u * u - w * w) * r12 - 2. * (15. * u * u + w * w) * c.f1) / d2;
    c.b12 = (2. * u * r21 + (3. * u * u - w * w) * r22 - 2. * (15. * u * u + w * w) * c.f2) / d2;
    c.b20 = (r01 - 3. * u * r02 - 24. * u * c.f0) / d2;
    c.b21 = (r11 - 3. * u * r12 - 24. * u * c.f1) / d2;
    c.b22 = (r21 - 3. * u * r22 - 24. * u * c.f2) / d2;
  }
  if (c.reflect) {
    c.f0 = conj(c.f0); c.f1 = -conj(c.f1); c.f2 = conj(c.f2);
    if (need_b) {
      c.b10 = conj(c.b10); c.b11 = -conj(c.b11); c.b12 = conj(c.b12);
      c.b20 = -conj(c.b20); c.b21 = conj(c.b21); c.b22 = -conj(c.b22);
    }
  }
  return c;
}

// field-hmc.h:82-123
Mat expiQ(const Mat& Q) {
  CH c = ch_coeffs(Q, false);
  if (c.zero) return Mat::unit();
  return Mat::unit() * c.f0 + Q * c.f1 + Q * Q * c.f2;
}

// field-multi.h:90-178
Mat DexpiQ(const Mat& Q, const Mat& DQ) {
  CH c = ch_coeffs(Q, true);
  if (c.zero) return Mat::zero();
  Mat B1 = Mat::unit() * c.b10 + Q * c.b11 + Q * Q * c.b12;
  Mat B2 = Mat::unit() * c.b20 + Q * c.b21 + Q * Q * c.b22;
  return B1 * tr(Q * DQ) + B2 * tr(Q * Q * DQ) + DQ * c.f1 + DQ * Q * c.f2 + Q * DQ * c.f2;
}

// field-multi.h:180-269
Mat compute_Lambda(const Mat& Q, const Mat& SigmaP, const Mat& U) {
  CH c = ch_coeffs(Q, true);
  if (c.zero) return Mat::zero();
  Mat B1 = Mat::unit() * c.b10 + Q * c.b11 + Q * Q * c.b12;
  Mat B2 = Mat::unit() * c.b20 + Q * c.b21 + Q * Q * c.b22;
  Mat Gamma = Q * tr(SigmaP * B1 * U) + Q * Q * tr(SigmaP * B2 * U) + U * SigmaP * c.f1 +
              Q * U * SigmaP * c.f2 + U * SigmaP * Q * c.f2;
  return hermitian_traceless(Gamma);
}

// ---------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), replacing qlat RngState.
void philox(const uint32_t in[4], const uint32_t k[2], uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3], k0 = k[0], k1 = k[1];
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

enum Stream : uint32_t { ST_MOMENTUM = 1, ST_HAAR = 2, ST_ACCEPT = 3, ST_DEMON = 4 };

void draw(uint64_t seed, uint32_t stream, uint64_t epoch, uint32_t index, uint32_t d, uint32_t out[4]) {
  uint32_t ctr[4] = {index, d, (uint32_t)epoch, stream | ((uint32_t)(epoch >> 32) << 8)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  philox(ctr, key, out);
}
// [0,1) with 53 random bits
inline double u01(uint32_t hi, uint32_t lo) {
  uint64_t x = ((uint64_t)hi << 32 | lo) >> 11;
  return (double)x * (1. / 9007199254740992.);
}
// (0,1]
inline double u01_open0(uint32_t hi, uint32_t lo) {
  uint64_t x = ((uint64_t)hi << 32 | lo) >> 11;
  return ((double)x + 1.) * (1. / 9007199254740992.);
}
// Box-Muller: one Philox block -> two N(0,1) doubles.
inline void gauss2(const uint32_t w[4], double& g0, double& g1) {
  double u1 = u01_open0(w[0], w[1]);
  double u2 = u01(w[2], w[3]);
  double r = std::sqrt(-2. * std::log(u1));
  double t = 6.283185307179586 * u2;
  g0 = r * std::cos(t);
  g1 = r * std::sin(t);
}

// ---------------------------------------------------------------- paths
// field-matrix.h:102-131. dir d in [0,4): multiply U_d(y), then y += e_d;
// d in [4,8): y -= e_{d-4}, multiply U_{d-4}(y)^+. Starts from the identity.
Mat path_product(const Field& f, const int* x, const int* dir, int n) {
  Mat mul = Mat::unit();
  int y[4] = {x[0], x[1], x[2], x[3]};
  for (int i = 0; i < n; ++i) {
    int d = dir[i];
    if (d < DIMN) {
      mul = mul * f.at(y, d);
      y[d]++;
    } else {
      y[d - DIMN]--;
      mul = mul * dagger(f.at(y, d - DIMN));
    }
  }
  return mul;
}

// field-hmc.h:203-242 (loop form at 226-240)
Mat staple_dagger(const Field& f, const int* x, int mu) {
  Mat s = Mat::zero();
  for (int nu = 0; nu < DIMN; ++nu) {
    if (nu == mu) continue;
    int d1[3] = {nu, mu, nu + DIMN};
    s += path_product(f, x, d1, 3);
    int d2[3] = {nu + DIMN, mu, nu};
    s += path_product(f, x, d2, 3);
  }
  return dagger(s);
}

// Undaggered staple sum (the C of hep-lat/0311018 without rho).
Mat staple_sum(const Field& f, const int* x, int mu) { return dagger(staple_dagger(f, x, mu)); }

// field-hmc.h:137-201
Mat rectangular_dagger(const Field& f, const int* x, int mu) {
  Mat r = Mat::zero();
  for (int nu = 0; nu < DIMN; ++nu) {
    if (mu == nu) continue;
    const int D = DIMN;
    int p[6][5] = {{nu, mu, mu, nu + D, mu + D},     {nu + D, mu, mu, nu, mu + D},
                   {nu, nu, mu, nu + D, nu + D},     {nu + D, nu + D, mu, nu, nu},
                   {mu + D, nu, mu, mu, nu + D},     {mu + D, nu + D, mu, mu, nu}};
    for (int k = 0; k < 6; ++k) r += path_product(f, x, p[k], 5);
  }
  return dagger(r);
}

// field-hmc.h:262-269
Mat extended_staple_dagger(const Field& f, const int* x, int mu, double c1) {
  double c0 = 1. - 8. * c1;
  Mat m0 = staple_dagger(f, x, mu);
  Mat m1 = rectangular_dagger(f, x, mu);
  return m0 * c0 + m1 * c1;
}

// field-staple.h:27-135: chair and twist index tables (defects kept as written).
std::vector<std::vector<int>> build_index(int mu, bool twist) {
  std::vector<std::vector<int>> out;
  const int D = DIMN;
  for (int nu = 0; nu < D; ++nu) {
    if (mu == nu) continue;
    for (int la = 0; la < D; ++la) {
      if (la == nu || la == mu) continue;
      int bases[4][4];
      if (!twist) {
        int b[4][4] = {{nu, la, la + D, nu + D}, {nu, la + D, la + D, nu},
                       {nu + D, la, la, nu + D}, {nu + D, la + D, la, nu}};
        std::memcpy(bases, b, sizeof b);
      } else {
        int b[4][4] = {{nu, la, nu + D, la + D}, {nu, la + D, nu + D, la},
                       {nu + D, la, nu, la + D}, {nu + D, la + D, nu, la}};
        std::memcpy(bases, b, sizeof b);
      }
      for (int k = 0; k < 4; ++k)
        for (int i = 1; i < 4; ++i) {
          std::vector<int> d(bases[k], bases[k] + 4);
          d.insert(d.begin() + i, mu);
          out.push_back(d);
        }
    }
  }
  return out;
}
const std::vector<std::vector<int>>& chair_index(int mu) {
  static std::vector<std::vector<int>> t[4] = {build_index(0, false), build_index(1, false),
                                               build_index(2, false), build_index(3, false)};
  return t[mu];
}
const std::vector<std::vector<int>>& twist_index(int mu) {
  static std::vector<std::vector<int>> t[4] = {build_index(0, true), build_index(1, true),
                                               build_index(2, true), build_index(3, true)};
  return t[mu];
}
// field-staple.h:141-165: returns the undaggered sum, as the reference does.
Mat chair_staple(const Field& f, const int* x, int mu) {
  Mat acc = Mat::zero();
  for (const auto& d : chair_index(mu)) acc = acc + path_product(f, x, d.data(), (int)d.size());
  return acc;
}
Mat twist_staple(const Field& f, const int* x, int mu) {
  Mat acc = Mat::zero();
  for (const auto& d : twist_index(mu)) acc = acc + path_product(f, x, d.data(), (int)d.size());
  return acc;
}

// field-hmc.h:309-337
int is_constrained(const int* x, int mu, int mag) {
  if (!mag) return 0;
  bool c = true;
  for (int i = 0; i < 4; ++i) {
    if (i == mu) continue;
    c = c && (x[i] % mag == 0);
  }
  if (c) {
    if (x[mu] % mag == mag - 1) return 100;
    if (x[mu] % mag == 0) return 10;
    return 1;
  }
  return 0;
}

// ---------------------------------------------------------------- reductions
// field-matrix.h:202-221
double get_plaq(const Field& f, const int* x) {
  double ret = 0.;
  for (int mu = 0; mu < DIMN; ++mu)
    for (int nu = 0; nu < mu; ++nu) {
      int xm[4] = {x[0], x[1], x[2], x[3]}; xm[mu]++;
      int xn[4] = {x[0], x[1], x[2], x[3]}; xn[nu]++;
      Mat m = f.at(x, mu) * f.at(xm, nu);
      Mat n = dagger(f.at(x, nu) * f.at(xn, mu));
      ret += retr(m * n);
    }
  return ret;
}

// field-matrix.h:249-297
double get_rectangular(const Field& f, const int* x) {
  double sum = 0.;
  for (int mu = 0; mu < DIMN; ++mu)
    for (int nu = 0; nu < mu; ++nu) {
      auto sh = [&](int a, int na, int b, int nb) {
        static thread_local int y[4];
        y[0] = x[0]; y[1] = x[1]; y[2] = x[2]; y[3] = x[3];
        y[a] += na;
        if (nb) y[b] += nb;
        return y;
      };
      int y1[4], y2[4], y3[4];
      // gx_1_0[mu][nu] = U_nu(x+mu); gx_1_1[.][nu] = U_nu(x+mu+nu); ...
      std::memcpy(y1, sh(mu, 1, 0, 0), sizeof y1);
      std::memcpy(y2, sh(mu, 1, nu, 1), sizeof y2);
      std::memcpy(y3, sh(nu, 1, 0, 0), sizeof y3);
      int y4[4]; std::memcpy(y4, sh(nu, 2, 0, 0), sizeof y4);
      int y5[4]; std::memcpy(y5, sh(mu, 2, 0, 0), sizeof y5);
      Mat m = dagger(f.at(x, mu) * f.at(y1, nu) * f.at(y2, nu));
      sum += retr(f.at(x, nu) * f.at(y3, nu) * f.at(y4, mu) * m);
      m = dagger(f.at(x, mu) * f.at(y1, mu) * f.at(y5, nu));
      sum += retr(f.at(x, nu) * f.at(y3, mu) * f.at(y2, mu) * m);
    }
  return sum;
}

template <class Fn>
double site_sum(const Geo& g, Fn fn) {
  std::vector<double> part(g.V);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < g.V; ++s) {
    int x[4];
    g.coord(s, x);
    part[s] = fn(x, s);
  }
  double acc = 0.;
  for (long s = 0; s < g.V; ++s) acc += part[s];
  return acc;
}

double total_plaq(const Field& f) {  // field-matrix.h:299-330
  return site_sum(*f.g, [&](const int* x, long) { return get_plaq(f, x); });
}
double total_rect(const Field& f) {  // field-matrix.h:354-385
  return site_sum(*f.g, [&](const int* x, long) { return get_rectangular(f, x); });
}
// field-hmc.h:596-629 (class-100 links carry no momentum)
double kinetic(const Field& p, int mag) {
  double s = site_sum(*p.g, [&](const int* x, long site) {
    double a = 0.;
    for (int mu = 0; mu < DIMN; ++mu) {
      if (is_constrained(x, mu, mag) == 100) continue;
      a += retr(p.at(site, mu) * p.at(site, mu));
    }
    return a;
  });
  return s / 2.;
}

// ---------------------------------------------------------------- stout
// Isotropic stout smearing (hep-lat/0311018 eqs. 1-3) built from get_Q
// (field-multi.h:32-39) with the 6-plaquette staple in place of the 1x2 one:
// Q = TA(C U^+) * (-i rho), U' = expiQ(Q) U.
struct StoutLevels {
  std::vector<std::vector<Mat>> U;    // U[0] thin ... U[n] smeared
  std::vector<std::vector<Mat>> Q;    // per level
  std::vector<std::vector<Mat>> eQ;   // expiQ(Q)
  std::vector<std::vector<Mat>> C;    // rho * staple sum
};

void stout_forward(const Geo& g, const std::vector<Mat>& U0, double rho, int n, StoutLevels& st) {
  st.U.assign(n + 1, {});
  st.Q.assign(n, {});
  st.eQ.assign(n, {});
  st.C.assign(n, {});
  st.U[0] = U0;
  for (int k = 0; k < n; ++k) {
    st.U[k + 1].resize(g.V * 4);
    st.Q[k].resize(g.V * 4);
    st.eQ[k].resize(g.V * 4);
    st.C[k].resize(g.V * 4);
    Field f{&g, st.U[k].data()};
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g.V; ++s) {
      int x[4];
      g.coord(s, x);
      for (int mu = 0; mu < DIMN; ++mu) {
        Mat stp = staple_sum(f, x, mu);
        Mat Om = trless_antiherm(stp * dagger(f.at(s, mu)));
        Mat Q = Om * cplx(0., -rho);
        Mat e = expiQ(Q);
        st.Q[k][s * 4 + mu] = Q;
        st.eQ[k][s * 4 + mu] = e;
        st.C[k][s * 4 + mu] = stp * rho;
        st.U[k + 1][s * 4 + mu] = e * f.at(s, mu);
      }
    }
  }
}

// Matrix derivative Z of the gauge action at the top level: for any variation
// dU of link (x,mu), dS = Re Tr(Z dU). S = -(beta/3) sum Re Tr(U Sigma^+).
void action_Z(const Geo& g, const std::vector<Mat>& U, const orc_action& a, std::vector<Mat>& Z) {
  Z.resize(g.V * 4);
  Field f{&g, const_cast<Mat*>(U.data())};
  const double fac = -a.beta / 3.;
#pragma omp parallel for schedule(static)
  for (long s = 0; s < g.V; ++s) {
    int x[4];
    g.coord(s, x);
    for (int mu = 0; mu < DIMN; ++mu) {
      Mat st = (a.type == 0) ? staple_dagger(f, x, mu) : extended_staple_dagger(f, x, mu, a.c1);
      Z[s * 4 + mu] = st * fac;
    }
  }
}

// One level of the stout backward chain (hep-lat/0311018 eqs. 72-75):
// Zk = Z' e^{iQ} + i C^+ Lambda + sum over nu of the six Lambda-inserted
// staple terms, with Y = -i rho U^+ Lambda. Lambda from compute_Lambda
// (field-multi.h:180-269) with SigmaP = Z'.
void stout_backward_level(const Geo& g, const std::vector<Mat>& U, const std::vector<Mat>& Q,
                          const std::vector<Mat>& eQ, const std::vector<Mat>& C, double rho,
                          const std::vector<Mat>& Zp, std::vector<Mat>& Z) {
  std::vector<Mat> Y(g.V * 4), Lam(g.V * 4);
#pragma omp parallel for schedule(static)
  for (long k = 0; k < g.V * 4; ++k) {
    Lam[k] = compute_Lambda(Q[k], Zp[k], U[k]);
    Y[k] = dagger(U[k]) * Lam[k] * cplx(0., -rho);
  }
  Z.resize(g.V * 4);
  Field f{&g, const_cast<Mat*>(U.data())};
  Field y{&g, Y.data()};
#pragma omp parallel for schedule(static)
  for (long s = 0; s < g.V; ++s) {
    int x[4];
    g.coord(s, x);
    for (int mu = 0; mu < DIMN; ++mu) {
      const long k = s * 4 + mu;
      Mat z = Zp[k] * eQ[k];
      z += dagger(C[k]) * Lam[k] * cplx(0., 1.);
      for (int nu = 0; nu < DIMN; ++nu) {
        if (nu == mu) continue;
        int xpm[4] = {x[0], x[1], x[2], x[3]}; xpm[mu]++;
        int xpn[4] = {x[0], x[1], x[2], x[3]}; xpn[nu]++;
        int xmn[4] = {x[0], x[1], x[2], x[3]}; xmn[nu]--;
        int xpmmn[4] = {x[0], x[1], x[2], x[3]}; xpmmn[mu]++; xpmmn[nu]--;
        z += f.at(xpm, nu) * dagger(f.at(xpn, mu)) * y.at(x, nu);                     // t1
        z += dagger(f.at(xpmmn, nu)) * dagger(f.at(xmn, mu)) * dagger(y.at(xmn, nu)); // t2
        z += dagger(y.at(xpm, nu)) * dagger(f.at(xpn, mu)) * dagger(f.at(x, nu));     // t3
        z += y.at(xpmmn, nu) * dagger(f.at(xmn, mu)) * f.at(xmn, nu);                 // t4
        z += dagger(f.at(xpmmn, nu)) * y.at(xmn, mu) * f.at(xmn, nu);                 // t5
        z += f.at(xpm, nu) * y.at(xpn, mu) * dagger(f.at(x, nu));                     // t6
      }
      Z[k] = z;
    }
  }
}

// Force from a matrix derivative Z with the constraint classes of
// field-hmc.h:353-378: class 0: F = -i TA(U Z); classes 1, 10:
// F = -i TA(U(x+mu) Z(x+mu) - Z(x) U(x)); class 100: 0. With
// Z = -(beta/3) Sigma^+ this is the reference's (i beta/3) TA(...).
void force_from_Z(const Geo& g, const std::vector<Mat>& U, const std::vector<Mat>& Z, int mag,
                  std::vector<Mat>& F) {
  F.resize(g.V * 4);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < g.V; ++s) {
    int x[4];
    g.coord(s, x);
    for (int mu = 0; mu < DIMN; ++mu) {
      const long k = s * 4 + mu;
      Mat t;
      switch (is_constrained(x, mu, mag)) {
        case 0: t = U[k] * Z[k]; break;
        case 1:
        case 10: {
          int y[4] = {x[0], x[1], x[2], x[3]}; y[mu]++;
          long ky = g.index(y) * 4 + mu;
          t = U[ky] * Z[ky] - Z[k] * U[k];
          break;
        }
        default: t = Mat::zero(); break;
      }
      F[k] = trless_antiherm(t) * cplx(0., -1.);
    }
  }
}

// field-hmc.h:339-425 (reference operation order for the thin-link force).
void get_force_thin(const Geo& g, const std::vector<Mat>& U, const orc_action& a, std::vector<Mat>& F) {
  F.resize(g.V * 4);
  Field f{&g, const_cast<Mat*>(U.data())};
#pragma omp parallel for schedule(static)
  for (long s = 0; s < g.V; ++s) {
    int x[4];
    g.coord(s, x);
    for (int mu = 0; mu < DIMN; ++mu) {
      auto stp = [&](const int* z) {
        return a.type == 0 ? staple_dagger(f, z, mu) : extended_staple_dagger(f, z, mu, a.c1);
      };
      Mat t;
      switch (is_constrained(x, mu, a.mag)) {
        case 0: t = f.at(s, mu) * stp(x); break;
        case 1:
        case 10: {
          int y[4] = {x[0], x[1], x[2], x[3]}; y[mu]++;
          Mat s1 = stp(x), s2 = stp(y);
          t = f.at(y, mu) * s2 - s1 * f.at(s, mu);
          break;
        }
        default: t = Mat::zero(); break;
      }
      F[s * 4 + mu] = trless_antiherm(t) * cplx(0., a.beta / 3.);
    }
  }
}

void get_force(const Geo& g, const std::vector<Mat>& U, const orc_action& a, std::vector<Mat>& F) {
  if (a.n_stout <= 0) { get_force_thin(g, U, a, F); return; }
  StoutLevels st;
  stout_forward(g, U, a.rho, a.n_stout, st);
  std::vector<Mat> Z, Zn;
  action_Z(g, st.U[a.n_stout], a, Z);
  for (int k = a.n_stout - 1; k >= 0; --k) {
    stout_backward_level(g, st.U[k], st.Q[k], st.eQ[k], st.C[k], a.rho, Z, Zn);
    Z.swap(Zn);
  }
  force_from_Z(g, U, Z, a.mag, F);
}

double gauge_action_of(const Geo& g, const std::vector<Mat>& U, const orc_action& a) {
  const std::vector<Mat>* top = &U;
  StoutLevels st;
  if (a.n_stout > 0) {
    stout_forward(g, U, a.rho, a.n_stout, st);
    top = &st.U[a.n_stout];
  }
  Field f{&g, const_cast<Mat*>(top->data())};
  // field-hmc.h:632-640
  if (a.type == 0) return -total_plaq(f) * a.beta / 3.;
  double p1 = -total_plaq(f);
  double p2 = -total_rect(f);
  return (p1 * (1. - 8. * a.c1) + p2 * a.c1) * a.beta / 3.;
}

// field-hmc.h:589-644
double hamiltonian(const Geo& g, const std::vector<Mat>& U, const std::vector<Mat>& P,
                   const orc_action& a, double* parts) {
  Field p{&g, const_cast<Mat*>(P.data())};
  double K = kinetic(p, a.mag);
  double S = gauge_action_of(g, U, a);
  if (parts) { parts[0] = K; parts[1] = S; }
  return K + S;
}

// ---------------------------------------------------------------- updates
// field-hmc.h:442-482
void evolve_gauge_field(const Geo& g, std::vector<Mat>& U, const std::vector<Mat>& P, double dt, int mag) {
#pragma omp parallel for schedule(static)
  for (long s = 0; s < g.V; ++s) {
    int x[4];
    g.coord(s, x);
    for (int mu = 0; mu < DIMN; ++mu) {
      const long k = s * 4 + mu;
      int y[4] = {x[0], x[1], x[2], x[3]}; y[mu]--;
      const long ky = g.index(y) * 4 + mu;
      switch (is_constrained(x, mu, mag)) {
        case 0: U[k] = expiQ(P[k] * dt) * U[k]; break;
        case 1: {
          Mat mL = expiQ(P[ky] * dt), mR = expiQ(P[k] * -dt);
          U[k] = mL * U[k] * mR;
          break;
        }
        case 10: U[k] = U[k] * expiQ(P[k] * -dt); break;
        case 100: U[k] = expiQ(P[ky] * dt) * U[k]; break;
      }
    }
  }
}

// field-hmc.h:427-440
void evolve_momentum(std::vector<Mat>& P, const std::vector<Mat>& F, double dt, double m) {
#pragma omp parallel for schedule(static)
  for (long k = 0; k < (long)P.size(); ++k) P[k] += F[k] * dt * m;
}

// field-matrix.h:83-100
double reunitarize(std::vector<Mat>& U) {
  std::vector<double> dev(U.size());
#pragma omp parallel for schedule(static)
  for (long k = 0; k < (long)U.size(); ++k) {
    Mat old = U[k];
    bool all_zero = true;
    for (int i = 0; i < 6; ++i) all_zero = all_zero && U[k].e[i].re == 0. && U[k].e[i].im == 0.;
    if (!all_zero) unitarize(U[k]);
    dev[k] = frob(U[k] - old);
  }
  double m = 0.;
  for (double d : dev) m = std::max(m, d);
  return m;
}

// field-hmc.h:525-567
double force_gradient(const Geo& g, std::vector<Mat>& U, std::vector<Mat>& P, const orc_action& a,
                      int steps, double dt, bool do_reunit) {
  const double alpha = (3. - std::sqrt(3.)) * dt / 6.;
  const double beta = dt / std::sqrt(3.);
  const double gamma = (2. - std::sqrt(3.)) * dt * dt / 12.;
  std::vector<Mat> Uaux, F;
  evolve_gauge_field(g, U, P, alpha, a.mag);
  for (int i = 0; i < steps; ++i) {
    for (int half = 0; half < 2; ++half) {
      get_force(g, U, a, F);
      Uaux = U;
      evolve_gauge_field(g, Uaux, F, gamma, a.mag);
      get_force(g, Uaux, a, F);
      evolve_momentum(P, F, 0.5 * dt, 1.);
      if (half == 0) evolve_gauge_field(g, U, P, beta, a.mag);
    }
    if (i < steps - 1) evolve_gauge_field(g, U, P, 2. * alpha, a.mag);
    else evolve_gauge_field(g, U, P, alpha, a.mag);
  }
  return do_reunit ? reunitarize(U) : 0.;
}

// field-hmc.h:569-587
void leap_frog(const Geo& g, std::vector<Mat>& U, std::vector<Mat>& P, const orc_action& a, int steps,
               double dt) {
  std::vector<Mat> F;
  evolve_gauge_field(g, U, P, dt / 2., a.mag);
  for (int i = 0; i < steps; ++i) {
    get_force(g, U, a, F);
    evolve_momentum(P, F, dt, 1.);
    if (i < steps - 1) evolve_gauge_field(g, U, P, dt, a.mag);
    else evolve_gauge_field(g, U, P, dt / 2., a.mag);
  }
}

// field-hmc.h:646-673 with the Philox momentum stream.
void init_momentum(const Geo& g, uint64_t seed, uint64_t traj, double m, std::vector<Mat>& P) {
  P.resize(g.V * 4);
  const double sig = std::sqrt(m);
  const auto& T = gens();
#pragma omp parallel for schedule(static)
  for (long k = 0; k < g.V * 4; ++k) {
    double gv[8];
    for (uint32_t d = 0; d < 4; ++d) {
      uint32_t w[4];
      draw(seed, ST_MOMENTUM, traj, (uint32_t)k, d, w);
      gauss2(w, gv[2 * d], gv[2 * d + 1]);
    }
    Mat t = Mat::zero();
    for (int a = 0; a < 8; ++a) t += T[a] * (gv[a] * sig);
    P[k] = t;
  }
}

// Haar SU(3) (Mezzadri 2007): Gram-Schmidt of a complex Gaussian matrix
// (positive diagonal of R), then the det phase divided out.
Mat haar_su3(uint64_t seed, uint32_t k) {
  Mat z;
  for (uint32_t d = 0; d < 9; ++d) {
    uint32_t w[4];
    draw(seed, ST_HAAR, 0, k, d, w);
    double g0, g1;
    gauss2(w, g0, g1);
    z.e[d] = cplx(g0, g1);
  }
  // rows: orthonormalise 0, 1, 2
  for (int r = 0; r < 3; ++r) {
    for (int q = 0; q < r; ++q) {
      cplx c;
      for (int j = 0; j < 3; ++j) c = c + conj(z(q, j)) * z(r, j);
      for (int j = 0; j < 3; ++j) z(r, j) = z(r, j) - c * z(q, j);
    }
    double n = 0.;
    for (int j = 0; j < 3; ++j) n += norm2(z(r, j));
    n = 1. / std::sqrt(n);
    for (int j = 0; j < 3; ++j) z(r, j) = z(r, j) * n;
  }
  cplx d = det(z);
  double phi = std::atan2(d.im, d.re);
  return z * expix(-phi / 3.);
}

// field-wloop.h:26-98
double wilson_loop(const Field& f, const int* x, const int* r, int T) {
  int dir[512];
  int n = 0;
  for (int i = 0; i < DIMN - 1; ++i)
    for (int j = 0; j < r[i]; ++j) dir[n++] = i;
  for (int i = 0; i < T; ++i) dir[n++] = DIMN - 1;
  for (int i = DIMN - 2; i >= 0; --i)
    for (int j = 0; j < r[i]; ++j) dir[n++] = i + DIMN;
  for (int i = 0; i < T; ++i) dir[n++] = DIMN - 1 + DIMN;
  return retr(path_product(f, x, dir, n));
}

double avg_wilson_loop(const Field& f, const int* r, int T) {
  std::vector<std::array<int, 3>> perm;
  std::array<int, 3> rl = {r[0], r[1], r[2]};
  std::sort(rl.begin(), rl.end());
  do { perm.push_back(rl); } while (std::next_permutation(rl.begin(), rl.end()));
  double s = site_sum(*f.g, [&](const int* x, long) {
    double a = 0.;
    for (const auto& p : perm) a += wilson_loop(f, x, p.data(), T);
    return a;
  });
  return s / (perm.size() * (double)f.g->V);
}

// ---------------------------------------------------------------- md:: coarse/fine HMC
// field-multi.h and field-multi-md.h. The coarse lattice has half the fine
// extents; coarse site xc sits at fine site x = 2 xc. rho, XI0 and md_alpha
// are parameters (the reference leaves rho and XI0 to the includer).
enum : uint32_t { ST_HEATBATH = 5 };

// field-multi.h:27-31
Mat get_U2(const Field& f, const int* x, int mu) {
  int y[4] = {x[0], x[1], x[2], x[3]};
  y[mu]++;
  return f.at(x, mu) * f.at(y, mu);
}

// field-hmc.h:244-260: the six 2x1 staples from x to x + 2 mu
Mat staple_rect(const Field& f, const int* x, int mu) {
  Mat s = Mat::zero();
  for (int nu = 0; nu < DIMN; ++nu) {
    if (nu == mu) continue;
    int d1[4] = {nu, mu, mu, nu + DIMN};
    s += path_product(f, x, d1, 4);
    int d2[4] = {nu + DIMN, mu, mu, nu};
    s += path_product(f, x, d2, 4);
  }
  return s;
}

// field-multi.h:33-39: Q = -i rho TA(stp U2^+)
Mat get_Q(const Field& f, const int* x, int mu, double rho) {
  Mat Omega = staple_rect(f, x, mu) * dagger(get_U2(f, x, mu));
  return trless_antiherm(Omega) * cplx(0., -rho);
}

// field-matrix.h:177-200; a direction of -1 multiplies by ins
Mat path_insertion(const Field& f, const int* x, const int* dir, int n, const Mat& ins) {
  Mat mul = Mat::unit();
  int y[4] = {x[0], x[1], x[2], x[3]};
  for (int i = 0; i < n; ++i) {
    int d = dir[i];
    if (d == -1) {
      mul = mul * ins;
    } else if (d < DIMN) {
      mul = mul * f.at(y, d);
      y[d]++;
    } else {
      y[d - DIMN]--;
      mul = mul * dagger(f.at(y, d - DIMN));
    }
  }
  return mul;
}

// field-matrix.h:138-175: d/dt of the path product with U' = i P U on
// every link, summed over the positions of the path
Mat path_leftD(const Field& g, const Field& m, const int* x, const int* dir, int n) {
  Mat add = Mat::zero();
  for (int d = 0; d < n; ++d) {
    int y[4] = {x[0], x[1], x[2], x[3]};
    Mat mul = Mat::unit();
    for (int i = 0; i < n; ++i) {
      int k = dir[i];
      if (k < DIMN) {
        if (i == d) mul = mul * m.at(y, k) * cplx(0., 1.) * g.at(y, k);
        else mul = mul * g.at(y, k);
        y[k]++;
      } else {
        y[k - DIMN]--;
        if (i == d) mul = mul * dagger(m.at(y, k - DIMN) * cplx(0., 1.) * g.at(y, k - DIMN));
        else mul = mul * dagger(g.at(y, k - DIMN));
      }
    }
    add += mul;
  }
  return add;
}

// field-multi.h:384-407: {type, nu} of fine link (x, mu)
void stout_type(const int* x, int mu, int out[2]) {
  int cnt[4], n = 0;
  for (int i = 0; i < DIMN; ++i)
    if (x[i] % 2 == 1) cnt[n++] = i;
  out[0] = 0; out[1] = 0;
  if (n == 0) { out[0] = 1; return; }
  if (n == 1) {
    if (cnt[0] == mu) out[0] = 3;
    else { out[0] = 2; out[1] = cnt[0]; }
    return;
  }
  if (n == 2) {
    if (cnt[0] == mu) { out[0] = 4; out[1] = cnt[1]; return; }
    if (cnt[1] == mu) { out[0] = 4; out[1] = cnt[0]; return; }
  }
}

struct Multi {
  const Geo& gf;
  const Geo& gc;
  double rho, xi0, alpha;
};

void half(const int* y, int* yc) { for (int i = 0; i < 4; ++i) yc[i] = y[i] / 2; }

// md::get_Fforce (field-multi-md.h:116-368): the thin gauge force without
// link classes, plus the stout term of every fine link by its stout_type.
void multi_fforce(const Multi& M, const std::vector<Mat>& UF, const std::vector<Mat>& UC,
                  const std::vector<double>& xi, const orc_action& a, std::vector<Mat>& F) {
  orc_action thin = a;
  thin.mag = 0;
  thin.n_stout = 0;
  get_force_thin(M.gf, UF, thin, F);
  Field f{&M.gf, const_cast<Mat*>(UF.data())}, c{&M.gc, const_cast<Mat*>(UC.data())};
  std::vector<Mat> Lam(M.gc.V * 4);
  // 164-176
#pragma omp parallel for schedule(static)
  for (long s = 0; s < M.gc.V; ++s) {
    int xc[4], x[4];
    M.gc.coord(s, xc);
    for (int i = 0; i < 4; ++i) x[i] = 2 * xc[i];
    for (int mu = 0; mu < DIMN; ++mu)
      Lam[s * 4 + mu] = compute_Lambda(get_Q(f, x, mu, M.rho), dagger(c.at(s, mu)), get_U2(f, x, mu));
  }
  Field cl{&M.gc, Lam.data()};
  const double rho = M.rho;
  auto LAM = [&](const int* y, int mu) { int yc[4]; half(y, yc); return cl.at(yc, mu); };
  auto XI = [&](const int* y, int mu) { int yc[4]; half(y, yc); return xi[M.gc.index(yc) * 4 + mu]; };
  auto CG = [&](const int* y, int mu) { int yc[4]; half(y, yc); return c.at(yc, mu); };
  const int D = DIMN;
#pragma omp parallel for schedule(static)
  for (long s = 0; s < M.gf.V; ++s) {
    int x[4];
    M.gf.coord(s, x);
    for (int mu = 0; mu < DIMN; ++mu) {
      Mat t = Mat::zero();
      int sx[4] = {x[0], x[1], x[2], x[3]};
      sx[mu]++;
      int tn[2];
      stout_type(x, mu, tn);
      int y[4];
      auto set = [&](int d0, int n0, int d1, int n1) {
        for (int i = 0; i < 4; ++i) y[i] = x[i];
        y[d0] += n0;
        y[d1] += n1;
      };
      switch (tn[0]) {
        case 1: {  // 196-252
          set(mu, 1, mu, 0);
          t += f.at(y, mu) * dagger(CG(x, mu)) * expiQ(get_Q(f, x, mu, rho)) * XI(x, mu);
          for (int nu = 0; nu < DIMN; ++nu) {
            if (nu == mu) continue;
            int p1[6] = {mu, nu, mu + D, mu + D, nu + D, -1};
            t += path_insertion(f, sx, p1, 6, LAM(x, mu)) * cplx(0., rho) * XI(x, mu);
            int p2[6] = {mu, nu + D, mu + D, mu + D, nu, -1};
            t += path_insertion(f, sx, p2, 6, LAM(x, mu)) * cplx(0., rho) * XI(x, mu);
            int p3[6] = {nu, nu, mu + D, nu + D, nu + D, -1};
            t += path_insertion(f, sx, p3, 6, LAM(x, nu)) * cplx(0., -rho) * XI(x, nu);
            set(nu, -2, nu, 0);
            int p4[6] = {nu + D, nu + D, mu + D, -1, nu, nu};
            t += path_insertion(f, sx, p4, 6, LAM(y, nu)) * cplx(0., rho) * XI(y, nu);
          }
          break;
        }
        case 2: {  // 254-278
          const int nu = tn[1];
          set(nu, -1, nu, 0);
          int p1[6] = {mu, nu + D, mu + D, mu + D, -1, nu};
          t += path_insertion(f, sx, p1, 6, LAM(y, mu)) * cplx(0., -rho) * XI(y, mu);
          set(nu, 1, nu, 0);
          int p2[6] = {mu, nu, mu + D, mu + D, -1, nu + D};
          t += path_insertion(f, sx, p2, 6, LAM(y, mu)) * cplx(0., -rho) * XI(y, mu);
          break;
        }
        case 3: {  // 279-332
          set(mu, -1, mu, 0);
          t += dagger(CG(y, mu)) * expiQ(get_Q(f, y, mu, rho)) * f.at(y, mu) * XI(y, mu);
          for (int nu = 0; nu < DIMN; ++nu) {
            if (nu == mu) continue;
            set(mu, -1, mu, 0);
            int p1[6] = {nu, mu + D, mu + D, nu + D, -1, mu};
            t += path_insertion(f, sx, p1, 6, LAM(y, mu)) * cplx(0., rho) * XI(y, mu);
            int p2[6] = {nu + D, mu + D, mu + D, nu, -1, mu};
            t += path_insertion(f, sx, p2, 6, LAM(y, mu)) * cplx(0., rho) * XI(y, mu);
            set(mu, 1, mu, 0);
            int p3[6] = {-1, nu, nu, mu + D, nu + D, nu + D};
            t += path_insertion(f, sx, p3, 6, LAM(y, nu)) * cplx(0., rho) * XI(y, nu);
            set(mu, 1, nu, -2);
            int p4[6] = {nu + D, nu + D, -1, mu + D, nu, nu};
            t += path_insertion(f, sx, p4, 6, LAM(y, nu)) * cplx(0., -rho) * XI(y, nu);
          }
          break;
        }
        case 4: {  // 333-357
          const int nu = tn[1];
          set(nu, -1, mu, -1);
          int p1[6] = {nu + D, mu + D, mu + D, -1, nu, mu};
          t += path_insertion(f, sx, p1, 6, LAM(y, mu)) * cplx(0., -rho) * XI(y, mu);
          set(nu, 1, mu, -1);
          int p2[6] = {nu, mu + D, mu + D, -1, nu + D, mu};
          t += path_insertion(f, sx, p2, 6, LAM(y, mu)) * cplx(0., -rho) * XI(y, mu);
          break;
        }
        default: break;
      }
      // 363-365
      Mat w = trless_antiherm(f.at(s, mu) * t);
      F[s * 4 + mu] += w * cplx(0., 1.);
    }
  }
}

// get_DQ (field-multi-md.h:370-396)
Mat get_DQ(const Field& g, const Field& p, const int* x, int mu, double rho) {
  Mat series = Mat::zero();
  for (int nu = 0; nu < DIMN; ++nu) {
    if (nu == mu) continue;
    int d1[6] = {nu, mu, mu, nu + DIMN, mu + DIMN, mu + DIMN};
    series += path_leftD(g, p, x, d1, 6);
    int d2[6] = {nu + DIMN, mu, mu, nu, mu + DIMN, mu + DIMN};
    series += path_leftD(g, p, x, d2, 6);
  }
  return trless_antiherm(series) * (I() * -rho);
}

// get_Cforce (field-multi-md.h:398-425): the coarse links' momentum, the
// velocity of Gb = e^{iQ} U2 along the fine momenta, scaled by alpha/(1+alpha)
void multi_cforce(const Multi& M, const std::vector<Mat>& UF, const std::vector<Mat>& PF, std::vector<Mat>& CF) {
  CF.resize(M.gc.V * 4);
  Field g{&M.gf, const_cast<Mat*>(UF.data())}, p{&M.gf, const_cast<Mat*>(PF.data())};
#pragma omp parallel for schedule(static)
  for (long s = 0; s < M.gc.V; ++s) {
    int xc[4], x[4];
    M.gc.coord(s, xc);
    for (int i = 0; i < 4; ++i) x[i] = 2 * xc[i];
    for (int mu = 0; mu < DIMN; ++mu) {
      int dir[2] = {mu, mu};
      Mat Q = get_Q(g, x, mu, M.rho);
      Mat U = get_U2(g, x, mu);
      Mat DQ = get_DQ(g, p, x, mu, M.rho);
      Mat DU = path_leftD(g, p, x, dir, 2);
      Mat t = (DexpiQ(Q, DQ) * U + expiQ(Q) * DU) * dagger(expiQ(Q) * U);
      CF[s * 4 + mu] = t * (-1. * I() * M.alpha / (1. + M.alpha));
    }
  }
}

// get_xi_energy (field-multi-md.h:30-72)
double xi_energy(const Multi& M, const std::vector<Mat>& UF, const std::vector<Mat>& UC,
                 const std::vector<double>& xi) {
  Field f{&M.gf, const_cast<Mat*>(UF.data())}, c{&M.gc, const_cast<Mat*>(UC.data())};
  std::vector<double> v(M.gc.V);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < M.gc.V; ++s) {
    int xc[4], x[4];
    M.gc.coord(s, xc);
    for (int i = 0; i < 4; ++i) x[i] = 2 * xc[i];
    double a = 0.;
    for (int mu = 0; mu < DIMN; ++mu) {
      Mat Gb = expiQ(get_Q(f, x, mu, M.rho)) * get_U2(f, x, mu);
      a += retr(dagger(c.at(s, mu)) * Gb) * xi[s * 4 + mu];
    }
    v[s] = a;
  }
  double sum = 0.;
  for (double a : v) sum += a;
  return -sum * (1. + M.alpha);
}

// get_Gb (field-multi-md.h:771-783)
void multi_gb(const Multi& M, const std::vector<Mat>& UF, std::vector<Mat>& Gb) {
  Gb.resize(M.gc.V * 4);
  Field f{&M.gf, const_cast<Mat*>(UF.data())};
#pragma omp parallel for schedule(static)
  for (long s = 0; s < M.gc.V; ++s) {
    int xc[4], x[4];
    M.gc.coord(s, xc);
    for (int i = 0; i < 4; ++i) x[i] = 2 * xc[i];
    for (int mu = 0; mu < DIMN; ++mu) Gb[s * 4 + mu] = expiQ(get_Q(f, x, mu, M.rho)) * get_U2(f, x, mu);
  }
}

// heatbath + simplest_metropolis (field-multi-md.h:748-769) with the
// Philox heat-bath stream keyed by (coarse link, trajectory): draws 0-1 give
// a0..a3 of random_su3_from_su2(0.9) (field-matrix.h:36-65), draw 2 the
// SU(2) subgroup and the Metropolis uniform.
void heatbath(const Multi& M, std::vector<Mat>& UC, const std::vector<Mat>& UF, uint64_t seed, uint64_t traj) {
  Field f{&M.gf, const_cast<Mat*>(UF.data())};
  const double coeff = (1. + M.alpha) * M.xi0;
#pragma omp parallel for schedule(static)
  for (long s = 0; s < M.gc.V; ++s) {
    int xc[4], x[4];
    M.gc.coord(s, xc);
    for (int i = 0; i < 4; ++i) x[i] = 2 * xc[i];
    for (int mu = 0; mu < DIMN; ++mu) {
      Mat env = expiQ(get_Q(f, x, mu, M.rho)) * get_U2(f, x, mu);
      uint32_t w0[4], w1[4], w2[4];
      const uint32_t link = (uint32_t)(s * 4 + mu);
      draw(seed, ST_HEATBATH, traj, link, 0, w0);
      draw(seed, ST_HEATBATH, traj, link, 1, w1);
      draw(seed, ST_HEATBATH, traj, link, 2, w2);
      double a0 = 0.9 + (1. - 0.9) * u01(w0[0], w0[1]);
      double a1 = -1. + 2. * u01(w0[2], w0[3]);
      double a2 = -1. + 2. * u01(w1[0], w1[1]);
      double a3 = -1. + 2. * u01(w1[2], w1[3]);
      double nrm = std::sqrt((1. - a0 * a0) / (a1 * a1 + a2 * a2 + a3 * a3));
      a1 *= nrm; a2 *= nrm; a3 *= nrm;
      Mat m = Mat::unit();
      switch (w2[0] % 3) {
        case 0: m(1, 1) = cplx(a0, a3); m(1, 2) = cplx(a2, a1); m(2, 1) = cplx(-a2, a1); m(2, 2) = cplx(a0, -a3); break;
        case 1: m(0, 0) = cplx(a0, a3); m(0, 2) = cplx(a2, a1); m(2, 0) = cplx(-a2, a1); m(2, 2) = cplx(a0, -a3); break;
        case 2: m(0, 0) = cplx(a0, a3); m(0, 1) = cplx(a2, a1); m(1, 0) = cplx(-a2, a1); m(1, 1) = cplx(a0, -a3); break;
      }
      Mat& var = UC[s * 4 + mu];
      Mat nv = var * m;
      double diff = coeff * retr(nv * dagger(env) - var * dagger(env));
      if (u01(w2[2], w2[3]) < std::exp(diff)) var = nv;
    }
  }
}

// force_gradient_integrator_zeta_T (field-multi-md.h:441-475)
void zeta_T(const Multi& M, std::vector<Mat>& UF, const std::vector<Mat>& PF, std::vector<Mat>& UC,
            std::vector<Mat>& Cm, double dt, int mag) {
  const double alpha = 1. / 6., beta = 2. / 3.;
  evolve_gauge_field(M.gf, UF, PF, alpha * dt, mag);
  multi_cforce(M, UF, PF, Cm);
  evolve_gauge_field(M.gc, UC, Cm, 0.5 * dt, mag);
  evolve_gauge_field(M.gf, UF, PF, beta * dt, mag);
  multi_cforce(M, UF, PF, Cm);
  evolve_gauge_field(M.gc, UC, Cm, 0.5 * dt, mag);
  evolve_gauge_field(M.gf, UF, PF, alpha * dt, mag);
}

// force_gradient_integrator_zeta_H (field-multi-md.h:477-597)
void zeta_H(const Multi& M, std::vector<Mat>& UF, std::vector<Mat>& PF, std::vector<Mat>& UC,
            const std::vector<double>& xi, const orc_action& a, int steps, double dt, double dev[2]) {
  const double alpha = (3. - std::sqrt(3.)) * dt / 6.;
  const double beta = dt / std::sqrt(3.);
  const double gamma = (2. - std::sqrt(3.)) * dt * dt / 12.;
  std::vector<Mat> Ff, FgA, CgA, Cm;
  zeta_T(M, UF, PF, UC, Cm, alpha, a.mag);
  for (int i = 0; i < steps; ++i) {
    for (int h = 0; h < 2; ++h) {
      multi_fforce(M, UF, UC, xi, a, Ff);
      FgA = UF;
      CgA = UC;
      zeta_T(M, FgA, Ff, CgA, Cm, gamma, a.mag);
      multi_fforce(M, FgA, CgA, xi, a, Ff);
      evolve_momentum(PF, Ff, 0.5 * dt, 1.);
      if (h == 0) zeta_T(M, UF, PF, UC, Cm, beta, a.mag);
    }
    zeta_T(M, UF, PF, UC, Cm, i < steps - 1 ? 2. * alpha : alpha, a.mag);
  }
  dev[0] = reunitarize(UF);
  dev[1] = reunitarize(UC);
}

// md::force_gradient_integrator (field-multi-md.h:599-661): fine fields only,
// the coarse links fixed. Reads its step size from dt (the reference's Farg.dt).
void multi_fg(const Multi& M, std::vector<Mat>& UF, std::vector<Mat>& PF, std::vector<Mat>& UC,
              const std::vector<double>& xi, const orc_action& a, int steps, double dt, double dev[2]) {
  const double alpha = (3. - std::sqrt(3.)) * dt / 6.;
  const double beta = dt / std::sqrt(3.);
  const double gamma = (2. - std::sqrt(3.)) * dt * dt / 12.;
  std::vector<Mat> Ff, FgA;
  evolve_gauge_field(M.gf, UF, PF, alpha, a.mag);
  for (int i = 0; i < steps; ++i) {
    for (int h = 0; h < 2; ++h) {
      multi_fforce(M, UF, UC, xi, a, Ff);
      FgA = UF;
      evolve_gauge_field(M.gf, FgA, Ff, gamma, a.mag);
      multi_fforce(M, FgA, UC, xi, a, Ff);
      evolve_momentum(PF, Ff, 0.5 * dt, 1.);
      if (h == 0) evolve_gauge_field(M.gf, UF, PF, beta, a.mag);
    }
    evolve_gauge_field(M.gf, UF, PF, i < steps - 1 ? 2. * alpha : alpha, a.mag);
  }
  dev[0] = reunitarize(UF);
  dev[1] = reunitarize(UC);
}

// nested_integrator (field-multi-md.h:663-695)
void multi_nested(const Multi& M, std::vector<Mat>& UF, std::vector<Mat>& PF, std::vector<Mat>& UC,
                  const std::vector<double>& xi, const orc_action& a, int steps, double dt, double dev[2]) {
  std::vector<Mat> Cm;
  multi_cforce(M, UF, PF, Cm);
  evolve_gauge_field(M.gc, UC, Cm, dt / 2., a.mag);
  for (int i = 0; i < steps; ++i) {
    multi_fg(M, UF, PF, UC, xi, a, 1, dt, dev);
    multi_cforce(M, UF, PF, Cm);
    evolve_gauge_field(M.gc, UC, Cm, i < steps - 1 ? dt : dt / 2., a.mag);
  }
}

// md::get_hamiltonian_multi (field-multi-md.h:74-114): parts = {kinetic
// (all links, field-multi.h:271-301), thin gauge action, S_xi}
double multi_hamiltonian(const Multi& M, const std::vector<Mat>& UF, const std::vector<Mat>& PF,
                         const std::vector<Mat>& UC, const std::vector<double>& xi, const orc_action& a,
                         double* parts) {
  Field p{&M.gf, const_cast<Mat*>(PF.data())};
  orc_action thin = a;
  thin.n_stout = 0;
  const double K = kinetic(p, 0);
  const double S = gauge_action_of(M.gf, UF, thin);
  const double X = xi_energy(M, UF, UC, xi);
  if (parts) { parts[0] = K; parts[1] = S; parts[2] = X; }
  return K + S + X;
}

// ---------------------------------------------------------------- local Metropolis
// algGCtrnHeatBath (alg_gchb.h:107-387): constrained local Metropolis on the
// Wilson action. CPS's per-site LRG is replaced by a Philox stream keyed by
// (link, iteration): hit h of the update of link (x, mu) takes draws 3h,
// 3h+1, 3h+2: a, b, c, d, select, die, each mapped to [-1, 1) as LRG with
// SetInterval(1, -1) does. A pair update draws from the stream of its first link.
enum : uint32_t { ST_LOCAL = 6 };

struct LocalDraws {
  double u[6];
};
LocalDraws local_draws(uint64_t seed, uint64_t iter, uint32_t link, int hit) {
  LocalDraws r;
  for (int k = 0; k < 3; ++k) {
    uint32_t w[4];
    draw(seed, ST_LOCAL, iter, link, (uint32_t)(3 * hit + k), w);
    r.u[2 * k] = -1. + 2. * u01(w[0], w[1]);
    r.u[2 * k + 1] = -1. + 2. * u01(w[2], w[3]);
  }
  return r;
}

// getRandomSU3BySubsu2 (alg_gchb.h:32-96) from the uniforms u[0..4]
Mat random_su3_by_subsu2(double small, const double* u) {
  double a = small * u[0];
  if (a > 0) a *= -1.;
  a += 1.;
  double b = u[1], c = u[2], d = u[3];
  const double norm = std::sqrt((1. - a * a) / (b * b + c * c + d * d));
  b *= norm; c *= norm; d *= norm;
  const double select = u[4];
  Mat m = Mat::zero();
  if (select < -1. / 3.) {
    m(0, 0) = cplx(a, b); m(0, 1) = cplx(c, d);
    m(1, 0) = cplx(-c, d); m(1, 1) = cplx(a, -b);
    m(2, 2) = cplx(1.);
  } else if (select > 1. / 3.) {
    m(0, 0) = cplx(1.);
    m(1, 1) = cplx(a, b); m(1, 2) = cplx(c, d);
    m(2, 1) = cplx(-c, d); m(2, 2) = cplx(a, -b);
  } else {
    m(0, 0) = cplx(a, b); m(0, 2) = cplx(c, d);
    m(1, 1) = cplx(1.);
    m(2, 0) = cplx(-c, d); m(2, 2) = cplx(a, -b);
  }
  return m;
}

struct LocalArgs {
  double beta, small;
  int nhits;
  uint64_t seed, iter;
  long acc[4];  // accept, reject, acceptCons, rejectCons
};

// metropolisKernel (alg_gchb.h:128-183)
void metropolis_kernel(Mat& U, const Mat& sigma, LocalArgs& A, uint32_t link) {
  const double fb = A.beta / 3.;
  for (int hit = 0; hit < A.nhits; ++hit) {
    LocalDraws r = local_draws(A.seed, A.iter, link, hit);
    Mat eps = random_su3_by_subsu2(A.small, r.u);
    const double oldS = -fb * retr(U * sigma);
    Mat t1 = U * eps;
    const double newS = -fb * retr(t1 * sigma);
    double die = r.u[5];
    if (die < 0.) die *= -1.;
    if (die < std::exp(oldS - newS)) { U = t1; A.acc[0]++; }
    else A.acc[1]++;
  }
}

// metropolisKernelPProd (alg_gchb.h:185-254): U1 eps and eps^+ U2 together
void metropolis_pair(Mat& U1, const Mat& s1, Mat& U2, const Mat& s2, LocalArgs& A, uint32_t link) {
  const double fb = A.beta / 3.;
  for (int hit = 0; hit < A.nhits; ++hit) {
    LocalDraws r = local_draws(A.seed, A.iter, link, hit);
    Mat eps = random_su3_by_subsu2(A.small, r.u);
    double oldS = -fb * retr(U1 * s1);
    oldS += -fb * retr(U2 * s2);
    Mat t1 = U1 * eps, t2 = dagger(eps) * U2;
    double newS = -fb * retr(t1 * s1);
    newS += -fb * retr(t2 * s2);
    double die = r.u[5];
    if (die < 0.) die *= -1.;
    if (die < std::exp(oldS - newS)) { U1 = t1; U2 = t2; A.acc[2]++; }
    else A.acc[3]++;
  }
}

// alg_gchb.h:117-127: the transverse coordinates are multiples of mag
bool gchb_constrained(const int* x, int mu, int mag) {
  bool c = true;
  for (int i = 0; i < 4; ++i)
    if (i != mu) c = c && (x[i] % mag == 0);
  return c;
}

// One link's update in run() (alg_gchb.h:327-368).
void local_update(Field& f, const int* x, int mu, int mag, LocalArgs& A) {
  const uint32_t link = (uint32_t)(f.g->index(x) * 4 + mu);
  if (mag > 0 && gchb_constrained(x, mu, mag) && x[mu] % mag != mag - 1) {
    int y[4] = {x[0], x[1], x[2], x[3]};
    y[mu]++;
    Mat s1 = staple_dagger(f, x, mu), s2 = staple_dagger(f, y, mu);
    metropolis_pair(f.at(x, mu), s1, f.at(y, mu), s2, A, link);
  } else {
    Mat s = staple_dagger(f, x, mu);
    metropolis_kernel(f.at(x, mu), s, A, link);
  }
}

}  // namespace

// ================================================================= C ABI
extern "C" {

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox(ctr, key, out);
}

void orc_generators(double* out) {
  for (int a = 0; a < 8; ++a) gens()[a].store(out + 18 * a);
}

void orc_expiQ(const double* Q, double* out, long n) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    Mat q; q.load(Q + 18 * i);
    expiQ(q).store(out + 18 * i);
  }
}

void orc_DexpiQ(const double* Q, const double* DQ, double* out, long n) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    Mat q, dq; q.load(Q + 18 * i); dq.load(DQ + 18 * i);
    DexpiQ(q, dq).store(out + 18 * i);
  }
}

void orc_compute_Lambda(const double* Q, const double* S, const double* U, double* out, long n) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    Mat q, s, u; q.load(Q + 18 * i); s.load(S + 18 * i); u.load(U + 18 * i);
    compute_Lambda(q, s, u).store(out + 18 * i);
  }
}

void orc_haar_links(const int* L, uint64_t seed, double* U) {
  Geo g(L);
#pragma omp parallel for schedule(static)
  for (long k = 0; k < g.V * 4; ++k) haar_su3(seed, (uint32_t)k).store(U + 18 * k);
}

void orc_unit_links(const int* L, double* U) {
  Geo g(L);
  for (long k = 0; k < g.V * 4; ++k) Mat::unit().store(U + 18 * k);
}

void orc_init_momentum(const int* L, uint64_t seed, uint64_t traj, double m, double* P) {
  Geo g(L);
  std::vector<Mat> p;
  init_momentum(g, seed, traj, m, p);
  store_field(p, P);
}

double orc_uniform(uint64_t seed, uint32_t stream, uint64_t index) {
  uint32_t w[4];
  draw(seed, stream, 0, (uint32_t)index, (uint32_t)(index >> 32), w);
  return u01(w[0], w[1]);
}

void orc_path_product(const int* L, const double* U, const int* x, const int* dir, int ndir, double* out) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U);
  Field f{&g, u.data()};
  path_product(f, x, dir, ndir).store(out);
}

#define ORC_SITE_FN(NAME, EXPR)                                              \
  void NAME(const int* L, const double* U, long site, int mu, double* out) { \
    Geo g(L);                                                                \
    std::vector<Mat> u = load_field(g, U);                                   \
    Field f{&g, u.data()};                                                   \
    int x[4];                                                                \
    g.coord(site, x);                                                        \
    (EXPR).store(out);                                                       \
  }
ORC_SITE_FN(orc_staple_dagger, staple_dagger(f, x, mu))
ORC_SITE_FN(orc_rectangular_dagger, rectangular_dagger(f, x, mu))
ORC_SITE_FN(orc_chair_staple_dagger, chair_staple(f, x, mu))
ORC_SITE_FN(orc_twist_staple_dagger, twist_staple(f, x, mu))

// The chair (twist = 0) or twist (twist = 1) path table of direction mu:
// writes up to max paths of 5 direction codes each and returns the count.
int orc_staple_index(int mu, int twist, int* out, int max) {
  const auto& t = twist ? twist_index(mu) : chair_index(mu);
  const int n = (int)t.size();
  for (int i = 0; i < n && i < max; ++i)
    for (int k = 0; k < 5; ++k) out[5 * i + k] = t[i][k];
  return n;
}

int orc_is_constrained(const int* x, int mu, int mag) { return is_constrained(x, mu, mag); }

double orc_total_plaq(const int* L, const double* U) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U);
  return total_plaq(Field{&g, u.data()});
}
double orc_total_rectangular(const int* L, const double* U) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U);
  return total_rect(Field{&g, u.data()});
}
double orc_avg_plaquette(const int* L, const double* U) {  // field-matrix.h:387-389
  Geo g(L);
  return orc_total_plaq(L, U) / (18. * g.V);
}
double orc_kinetic(const int* L, const double* P, int mag) {
  Geo g(L);
  std::vector<Mat> p = load_field(g, P);
  return kinetic(Field{&g, p.data()}, mag);
}
double orc_gauge_action(const int* L, const double* U, const orc_action* a) {
  Geo g(L);
  return gauge_action_of(g, load_field(g, U), *a);
}
double orc_hamiltonian(const int* L, const double* U, const double* P, const orc_action* a, double* parts) {
  Geo g(L);
  return hamiltonian(g, load_field(g, U), load_field(g, P), *a, parts);
}
double orc_reunitarize(const int* L, double* U) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U);
  double d = reunitarize(u);
  store_field(u, U);
  return d;
}

void orc_get_force(const int* L, const double* U, double* F, const orc_action* a) {
  Geo g(L);
  std::vector<Mat> f;
  get_force(g, load_field(g, U), *a, f);
  store_field(f, F);
}

void orc_evolve_gauge_field(const int* L, double* U, const double* P, double dt, int mag) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U);
  evolve_gauge_field(g, u, load_field(g, P), dt, mag);
  store_field(u, U);
}

void orc_evolve_momentum(const int* L, double* P, const double* F, double dt, double m) {
  Geo g(L);
  std::vector<Mat> p = load_field(g, P);
  evolve_momentum(p, load_field(g, F), dt, m);
  store_field(p, P);
}

void orc_stout_smear(const int* L, const double* U, double rho, int nlev, double* out) {
  Geo g(L);
  StoutLevels st;
  stout_forward(g, load_field(g, U), rho, nlev, st);
  store_field(st.U[nlev], out);
}

void orc_stout_Q(const int* L, const double* U, double rho, double* Q) {
  Geo g(L);
  StoutLevels st;
  stout_forward(g, load_field(g, U), rho, 1, st);
  store_field(st.Q[0], Q);
}

double orc_force_gradient_integrator(const int* L, double* U, double* P, const orc_action* a, int steps,
                                     double dt, int do_reunit) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U), p = load_field(g, P);
  double d = force_gradient(g, u, p, *a, steps, dt, do_reunit != 0);
  store_field(u, U);
  store_field(p, P);
  return d;
}

void orc_leap_frog_integrator(const int* L, double* U, double* P, const orc_action* a, int steps, double dt) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U), p = load_field(g, P);
  leap_frog(g, u, p, *a, steps, dt);
  store_field(u, U);
  store_field(p, P);
}

double orc_avg_wilson_loop(const int* L, const double* U, const int* r, int T) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U);
  return avg_wilson_loop(Field{&g, u.data()}, r, T);
}

// field-hmc.h:770-904 (run_chmc) without file output; trajectory index i
// keys the momentum stream and the die roll. Accept modes: 0 Metropolis
// (field-hmc.h:818-843), 1 demon window on (plaquette, rectangle) action sums
// relative to the chain's initial values (field-demon.h:152-184, commented
// in the reference), 2 both.
int orc_run_hmc(const int* L, double* U, const orc_action* a, const orc_hmc* h, int first_traj, int ntraj,
                double* ds, orc_traj* out) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U), p, backup;
  int naccept = 0;
  for (int t = 0; t < ntraj; ++t) {
    const int i = first_traj + t;
    init_momentum(g, h->seed, (uint64_t)i, 1., p);
    double parts_old[2], parts_new[2];
    double h_old = hamiltonian(g, u, p, *a, parts_old);
    backup = u;
    double dev = 0.;
    if (h->integrator == 0) dev = force_gradient(g, u, p, *a, h->trajectory_length, h->dt, h->reunitarize != 0);
    else leap_frog(g, u, p, *a, h->trajectory_length, h->dt);
    double h_new = hamiltonian(g, u, p, *a, parts_new);
    double die = orc_uniform(h->seed, ST_ACCEPT, (uint64_t)i);
    double prob = std::exp(h_old - h_new);
    bool acc = die < prob;
    // demon window (plaquette and rectangle sums of the top-level field)
    double Sp = 0., Sr = 0.;
    if (h->accept_mode != 0) {
      StoutLevels st;
      const std::vector<Mat>* top = &u;
      if (a->n_stout > 0) { stout_forward(g, u, a->rho, a->n_stout, st); top = &st.U[a->n_stout]; }
      Field f{&g, const_cast<Mat*>(top->data())};
      Sp = -total_plaq(f);
      Sr = -total_rect(f);
      bool win = (ds[0] - Sp < h->demon_emax) && (ds[0] - Sp > -h->demon_emax) &&
                 (ds[1] - Sr < h->demon_emax) && (ds[1] - Sr > -h->demon_emax);
      acc = (h->accept_mode == 1) ? win : (acc && win);
    }
    const bool bad = std::isnan(h_new) || std::isnan(h_old);
    if (i < h->num_forced_accept && !bad) acc = true;
    if (bad) acc = false;
    if (acc) {
      if (h->accept_mode != 0) { ds[2] = ds[0] - Sp; ds[3] = ds[1] - Sr; }
      ++naccept;
    } else {
      u = backup;
    }
    if (out) {
      orc_traj& r = out[t];
      r.h_old = h_old; r.h_new = h_new; r.dh = h_new - h_old; r.die_roll = die;
      r.accept = acc ? 1 : 0;
      Field f{&g, u.data()};
      r.avg_plaq = total_plaq(f) / (18. * g.V);
      r.s_gauge = acc ? parts_new[1] : parts_old[1];
      r.demon_e[0] = ds[2]; r.demon_e[1] = ds[3];
      r.reunit_dev = dev;
    }
  }
  store_field(u, U);
  return naccept;
}

// field-demon.h:25-126: sequential demon updates over sites in index order,
// 2 hits per link with random_su3_from_su2(0.9) (field-matrix.h:36-65);
// Philox demon stream keyed by (link, hit, sweep).
void orc_demon_microcanonical(const int* L, double* U, int nsweeps, double E_max, uint64_t seed, double* E,
                              long* acc_rej) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U);
  Field f{&g, u.data()};
  long acc = 0, rej = 0;
  for (int sw = 0; sw < nsweeps; ++sw) {
    for (long s = 0; s < g.V; ++s) {
      int x[4];
      g.coord(s, x);
      for (int mu = 0; mu < DIMN; ++mu) {
        Mat& v = f.at(s, mu);
        Mat P = staple_dagger(f, x, mu), R = rectangular_dagger(f, x, mu);
        Mat C = chair_staple(f, x, mu), T = twist_staple(f, x, mu);
        for (int n = 0; n < 2; ++n) {
          uint32_t w0[4], w1[4], w2[4];
          const uint32_t link = (uint32_t)(s * 4 + mu);
          draw(seed, ST_DEMON, (uint64_t)sw, link, 3 * n, w0);
          draw(seed, ST_DEMON, (uint64_t)sw, link, 3 * n + 1, w1);
          draw(seed, ST_DEMON, (uint64_t)sw, link, 3 * n + 2, w2);
          double a0 = 0.9 + (1. - 0.9) * u01(w0[0], w0[1]);
          double a1 = -1. + 2. * u01(w0[2], w0[3]);
          double a2 = -1. + 2. * u01(w1[0], w1[1]);
          double a3 = -1. + 2. * u01(w1[2], w1[3]);
          double nrm = std::sqrt((1. - a0 * a0) / (a1 * a1 + a2 * a2 + a3 * a3));
          a1 *= nrm; a2 *= nrm; a3 *= nrm;
          Mat m = Mat::unit();
          switch (w2[0] % 3) {
            case 0: m(1, 1) = cplx(a0, a3); m(1, 2) = cplx(a2, a1); m(2, 1) = cplx(-a2, a1); m(2, 2) = cplx(a0, -a3); break;
            case 1: m(0, 0) = cplx(a0, a3); m(0, 2) = cplx(a2, a1); m(2, 0) = cplx(-a2, a1); m(2, 2) = cplx(a0, -a3); break;
            case 2: m(0, 0) = cplx(a0, a3); m(0, 1) = cplx(a2, a1); m(1, 0) = cplx(-a2, a1); m(1, 1) = cplx(a0, -a3); break;
          }
          double dS[4];
          Mat mv = m * v;
          dS[0] = -retr(mv * P) + retr(v * P);
          dS[1] = -retr(mv * R) + retr(v * R);
          dS[2] = -retr(mv * C) + retr(v * C);
          dS[3] = -retr(mv * T) + retr(v * T);
          bool ok = true;
          for (int q = 0; q < 4; ++q) ok = ok && (E[q] - dS[q] > -E_max) && (E[q] - dS[q] < E_max);
          if (ok) {
            v = mv;
            for (int q = 0; q < 4; ++q) E[q] -= dS[q];
            ++acc;
          } else {
            ++rej;
          }
        }
      }
    }
  }
  store_field(u, U);
  acc_rej[0] = acc;
  acc_rej[1] = rej;
}

// ------------------------------------------------- mag-blocked lattices
// Coarse link = ordered product of the mag fine links from mag*xC along mu:
// the coarse-field initialisation of double_multigrid (field-hmc.h:1068-1078).
void orc_coarsen(const int* Lf, const double* Uf, int mag, double* Uc) {
  Geo gf(Lf);
  int Lc[4];
  for (int i = 0; i < 4; ++i) Lc[i] = Lf[i] / mag;
  Geo gc(Lc);
  std::vector<Mat> u = load_field(gf, Uf), c(gc.V * 4);
  Field f{&gf, u.data()};
  std::vector<int> dir(mag);
  for (long s = 0; s < gc.V; ++s) {
    int xc[4], x[4];
    gc.coord(s, xc);
    for (int i = 0; i < 4; ++i) x[i] = mag * xc[i];
    for (int mu = 0; mu < 4; ++mu) {
      std::fill(dir.begin(), dir.end(), mu);
      c[s * 4 + mu] = path_product(f, x, dir.data(), mag);
    }
  }
  store_field(c, Uc);
}

// update_constrain (field-hmc.h:988-1009): the first fine link of every
// coarse link becomes U_c(xC, mu) [U(x+mu) ... U(x+(mag-1)mu)]^+, so the
// product of the mag fine links equals the coarse link.
void orc_update_constrain(const int* Lf, double* Uf, const double* Uc, int mag) {
  Geo gf(Lf);
  int Lc[4];
  for (int i = 0; i < 4; ++i) Lc[i] = Lf[i] / mag;
  Geo gc(Lc);
  std::vector<Mat> u = load_field(gf, Uf), c = load_field(gc, Uc);
  Field f{&gf, u.data()};
  std::vector<int> dir(mag > 1 ? mag - 1 : 1);
  for (long s = 0; s < gc.V; ++s) {
    int xc[4], x[4];
    gc.coord(s, xc);
    for (int i = 0; i < 4; ++i) x[i] = mag * xc[i];
    for (int mu = 0; mu < 4; ++mu) {
      int xp[4] = {x[0], x[1], x[2], x[3]};
      xp[mu]++;
      std::fill(dir.begin(), dir.end(), mu);
      const Mat m = path_product(f, xp, dir.data(), mag - 1);
      f.at(x, mu) = c[s * 4 + mu] * dagger(m);
    }
  }
  store_field(u, Uf);
}

// check_constrained_plaquette (field-matrix.h:407-443): Re Tr / 3 of the
// mag x mag loops at the coarse sites, all planes nu < mu, averaged.
double orc_constrained_plaquette(const int* L, const double* U, int mag) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U);
  Field f{&g, u.data()};
  long count = 0;
  double sum = 0.;
  std::vector<int> dir(4 * mag);
  for (int x0 = 0; x0 < L[0]; x0 += mag)
    for (int x1 = 0; x1 < L[1]; x1 += mag)
      for (int x2 = 0; x2 < L[2]; x2 += mag)
        for (int x3 = 0; x3 < L[3]; x3 += mag) {
          const int x[4] = {x0, x1, x2, x3};
          for (int mu = 0; mu < 4; ++mu)
            for (int nu = 0; nu < mu; ++nu) {
              for (int i = 0; i < mag; ++i) {
                dir[i] = mu;
                dir[mag + i] = nu;
                dir[2 * mag + i] = mu + DIMN;
                dir[3 * mag + i] = nu + DIMN;
              }
              sum += retr(path_product(f, x, dir.data(), 4 * mag));
              ++count;
            }
        }
  return sum / (3. * count);
}

// lattice_expansion (expansion/main.C:19-62) without the file I/O: unit links
// everywhere, the coarse link U(x, mu) at fine site mag*x.
void orc_expand(const int* Lc, const double* Uc, int mag, double* Uf) {
  Geo gc(Lc);
  int Lf[4];
  for (int i = 0; i < 4; ++i) Lf[i] = Lc[i] * mag;
  Geo gf(Lf);
  std::vector<Mat> c = load_field(gc, Uc), u(gf.V * 4, Mat::unit());
  Field f{&gf, u.data()};
  for (long s = 0; s < gc.V; ++s) {
    int xc[4], x[4];
    gc.coord(s, xc);
    for (int i = 0; i < 4; ++i) x[i] = mag * xc[i];
    for (int mu = 0; mu < 4; ++mu) f.at(x, mu) = c[s * 4 + mu];
  }
  store_field(u, Uf);
}

// ------------------------------------------------- md:: coarse/fine HMC
// Fine extents Lf; the coarse field has extents Lf/2. xi: [coarse site][mu].
static Multi make_multi(const Geo& gf, const Geo& gc, const orc_multi* mp) {
  return Multi{gf, gc, mp->rho, mp->xi0, mp->md_alpha};
}
static void coarse_dims(const int* Lf, int* Lc) { for (int i = 0; i < 4; ++i) Lc[i] = Lf[i] / 2; }

#define MULTI_GEO                \
  Geo gf(Lf);                    \
  int Lc_[4];                    \
  coarse_dims(Lf, Lc_);          \
  Geo gc(Lc_);                   \
  Multi M = make_multi(gf, gc, mp)

static std::vector<double> load_xi(const Geo& gc, const double* xi) { return std::vector<double>(xi, xi + gc.V * 4); }

double orc_xi_energy(const int* Lf, const double* UF, const double* UC, const double* xi, const orc_multi* mp) {
  MULTI_GEO;
  return xi_energy(M, load_field(gf, UF), load_field(gc, UC), load_xi(gc, xi));
}

void orc_multi_gb(const int* Lf, const double* UF, const orc_multi* mp, double* Gb) {
  MULTI_GEO;
  std::vector<Mat> g;
  multi_gb(M, load_field(gf, UF), g);
  store_field(g, Gb);
}

void orc_multi_Q(const int* Lf, const double* UF, const orc_multi* mp, double* Q) {
  MULTI_GEO;
  std::vector<Mat> u = load_field(gf, UF), q(gc.V * 4);
  Field f{&gf, u.data()};
  for (long s = 0; s < gc.V; ++s) {
    int xc[4], x[4];
    gc.coord(s, xc);
    for (int i = 0; i < 4; ++i) x[i] = 2 * xc[i];
    for (int mu = 0; mu < 4; ++mu) q[s * 4 + mu] = get_Q(f, x, mu, M.rho);
  }
  store_field(q, Q);
}

void orc_multi_fforce(const int* Lf, const double* UF, const double* UC, const double* xi, const orc_action* a,
                      const orc_multi* mp, double* F) {
  MULTI_GEO;
  std::vector<Mat> f;
  multi_fforce(M, load_field(gf, UF), load_field(gc, UC), load_xi(gc, xi), *a, f);
  store_field(f, F);
}

void orc_multi_cforce(const int* Lf, const double* UF, const double* PF, const orc_multi* mp, double* CF) {
  MULTI_GEO;
  std::vector<Mat> cf;
  multi_cforce(M, load_field(gf, UF), load_field(gf, PF), cf);
  store_field(cf, CF);
}

void orc_multi_DQ(const int* Lf, const double* UF, const double* PF, const orc_multi* mp, double* DQ) {
  MULTI_GEO;
  std::vector<Mat> u = load_field(gf, UF), p = load_field(gf, PF), d(gc.V * 4);
  Field g{&gf, u.data()}, m{&gf, p.data()};
  for (long s = 0; s < gc.V; ++s) {
    int xc[4], x[4];
    gc.coord(s, xc);
    for (int i = 0; i < 4; ++i) x[i] = 2 * xc[i];
    for (int mu = 0; mu < 4; ++mu) d[s * 4 + mu] = get_DQ(g, m, x, mu, M.rho);
  }
  store_field(d, DQ);
}

void orc_multi_heatbath(const int* Lf, double* UC, const double* UF, const orc_multi* mp, uint64_t seed,
                        uint64_t traj) {
  MULTI_GEO;
  std::vector<Mat> c = load_field(gc, UC);
  heatbath(M, c, load_field(gf, UF), seed, traj);
  store_field(c, UC);
}

void orc_multi_zeta_T(const int* Lf, double* UF, const double* PF, double* UC, double dt, int mag,
                      const orc_multi* mp) {
  MULTI_GEO;
  std::vector<Mat> u = load_field(gf, UF), c = load_field(gc, UC), cm;
  zeta_T(M, u, load_field(gf, PF), c, cm, dt, mag);
  store_field(u, UF);
  store_field(c, UC);
}

// which: 0 zeta_H, 1 md::force_gradient_integrator, 2 nested_integrator
void orc_multi_integrate(const int* Lf, double* UF, double* PF, double* UC, const double* xi, const orc_action* a,
                         const orc_multi* mp, int which, int steps, double dt, double* dev) {
  MULTI_GEO;
  std::vector<Mat> u = load_field(gf, UF), p = load_field(gf, PF), c = load_field(gc, UC);
  std::vector<double> x = load_xi(gc, xi);
  double d[2] = {0., 0.};
  if (which == 0) zeta_H(M, u, p, c, x, *a, steps, dt, d);
  else if (which == 1) multi_fg(M, u, p, c, x, *a, steps, dt, d);
  else multi_nested(M, u, p, c, x, *a, steps, dt, d);
  store_field(u, UF);
  store_field(p, PF);
  store_field(c, UC);
  if (dev) { dev[0] = d[0]; dev[1] = d[1]; }
}

double orc_multi_hamiltonian(const int* Lf, const double* UF, const double* PF, const double* UC, const double* xi,
                             const orc_action* a, const orc_multi* mp, double* parts) {
  MULTI_GEO;
  return multi_hamiltonian(M, load_field(gf, UF), load_field(gf, PF), load_field(gc, UC), load_xi(gc, xi), *a,
                           parts);
}

// md::run_hmc_multi (field-multi-md.h:785-988) without the file output:
// per trajectory init_momentum, heatbath of the coarse links, init_xi
// (xi = XI0), H, zeta_H, H, Metropolis with forced accepts. A reject
// restores both fields to the last accepted state (which also drops the
// trajectory's heat-bath update, as in the reference).
int orc_run_hmc_multi(const int* Lf, double* UF, double* UC, const orc_action* a, const orc_hmc* h,
                      const orc_multi* mp, int first_traj, int ntraj, orc_multi_traj* out) {
  MULTI_GEO;
  std::vector<Mat> u = load_field(gf, UF), c = load_field(gc, UC), p;
  std::vector<Mat> u_ext = u, c_ext = c;
  std::vector<double> xi(gc.V * 4);
  int naccept = 0;
  for (int t = 0; t < ntraj; ++t) {
    const int i = first_traj + t;
    init_momentum(gf, h->seed, (uint64_t)i, 1., p);
    heatbath(M, c, u, h->seed, (uint64_t)i);
    std::fill(xi.begin(), xi.end(), mp->xi0);
    double po[3], pn[3], dev[2];
    const double h_old = multi_hamiltonian(M, u, p, c, xi, *a, po);
    zeta_H(M, u, p, c, xi, *a, h->trajectory_length, h->dt, dev);
    const double h_new = multi_hamiltonian(M, u, p, c, xi, *a, pn);
    const double die = orc_uniform(h->seed, ST_ACCEPT, (uint64_t)i);
    const double prob = std::exp(h_old - h_new);
    bool acc = die < prob;
    const bool bad = std::isnan(h_new) || std::isnan(h_old);
    if (i < h->num_forced_accept && !bad) acc = true;
    if (bad) acc = false;
    if (acc) {
      u_ext = u;
      c_ext = c;
      ++naccept;
    } else {
      u = u_ext;
      c = c_ext;
    }
    if (out) {
      orc_multi_traj& r = out[t];
      r.h_old = h_old; r.h_new = h_new; r.dh = h_new - h_old; r.die_roll = die;
      r.accept = acc ? 1 : 0;
      Field ff{&gf, u.data()}, cf{&gc, c.data()};
      r.plaq_fine = total_plaq(ff) / (18. * gf.V);
      r.plaq_coarse = total_plaq(cf) / (18. * gc.V);
      for (int k = 0; k < 3; ++k) { r.parts_old[k] = po[k]; r.parts_new[k] = pn[k]; }
      r.reunit_dev[0] = dev[0]; r.reunit_dev[1] = dev[1];
    }
  }
  store_field(u, UF);
  store_field(c, UC);
  return naccept;
}

// algGCtrnHeatBath::run (alg_gchb.h:290-387) on the Wilson action: numIter
// iterations; per iteration every direction, every site in the reference's
// order (order 0; x0 fastest within mu) or in 16 checkerboard classes (order
// 1: mu, then x_mu parity, then the parity of the transverse coordinate sum;
// links of one class share no staple, so a class's updates commute), then
// reunitarize and record the average and constrained plaquettes. mag = 0
// updates every link alone. acc = {accept, reject, acceptCons, rejectCons}.
void orc_local_metropolis(const int* L, double* U, double beta, int mag, int num_iter, int nhits, double small,
                          uint64_t seed, int order, long* acc, double* plaq, double* cons_plaq) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U);
  Field f{&g, u.data()};
  LocalArgs A{beta, small, nhits, seed, 0, {0, 0, 0, 0}};
  for (int it = 0; it < num_iter; ++it) {
    A.iter = (uint64_t)it;
    for (int mu = 0; mu < 4; ++mu) {
      if (order == 0) {
        for (long s = 0; s < g.V; ++s) {
          int x[4];
          g.coord(s, x);
          local_update(f, x, mu, mag, A);
        }
      } else {
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b)
            for (long s = 0; s < g.V; ++s) {
              int x[4];
              g.coord(s, x);
              int tr = 0;
              for (int i = 0; i < 4; ++i)
                if (i != mu) tr += x[i];
              if (x[mu] % 2 == a && tr % 2 == b) local_update(f, x, mu, mag, A);
            }
      }
    }
    reunitarize(u);
    if (plaq) plaq[it] = total_plaq(f) / (18. * g.V);
    if (cons_plaq) {
      store_field(u, U);
      cons_plaq[it] = mag > 0 ? orc_constrained_plaquette(L, U, mag) : 0.;
    }
  }
  store_field(u, U);
  for (int k = 0; k < 4; ++k) acc[k] = A.acc[k];
}

// One link's metropolisKernel hits with a fixed staple sum (test hook).
void orc_metropolis_link(double* U, const double* sigma, double beta, int nhits, double small, uint64_t seed,
                         uint64_t iter, uint32_t link, long* acc) {
  Mat u, s;
  u.load(U);
  s.load(sigma);
  LocalArgs A{beta, small, nhits, seed, iter, {0, 0, 0, 0}};
  metropolis_kernel(u, s, A, link);
  u.store(U);
  acc[0] += A.acc[0];
  acc[1] += A.acc[1];
}

// Checkerboard demon: demon_microcanonical (field-demon.h:25-126) with a
// demon per link. Each link carries its own four energies E[(s*4+mu)*4+k]
// (plaquette, rectangle, chair, twist); an accepted hit moves the link's
// action change into them, within (-E_max, E_max) per demon. Per sweep and
// direction mu the links are visited in 128 classes (x_mu mod 2, then x_nu
// mod 4 for the transverse nu in increasing order) and in site order within a
// class; members of one class share no staple path, so their updates commute.
// Same proposals and Philox stream as the sequential demon.
void orc_demon_checkerboard(const int* L, double* U, int nsweeps, double E_max, uint64_t seed, double* Elink,
                            long* acc_rej) {
  Geo g(L);
  std::vector<Mat> u = load_field(g, U);
  Field f{&g, u.data()};
  long acc = 0, rej = 0;
  for (int sw = 0; sw < nsweeps; ++sw)
    for (int mu = 0; mu < DIMN; ++mu)
      for (int cls = 0; cls < 128; ++cls)
        for (long s = 0; s < g.V; ++s) {
          int x[4];
          g.coord(s, x);
          int c = x[mu] % 2, sh = 1;
          for (int i = 0; i < 4; ++i)
            if (i != mu) {
              c += (x[i] % 4) * 2 * sh;
              sh *= 4;
            }
          if (c != cls) continue;
          Mat& v = f.at(s, mu);
          Mat X[4] = {staple_dagger(f, x, mu), rectangular_dagger(f, x, mu), chair_staple(f, x, mu),
                      twist_staple(f, x, mu)};
          double* E = Elink + (s * 4 + mu) * 4;
          for (int n = 0; n < 2; ++n) {
            uint32_t w0[4], w1[4], w2[4];
            const uint32_t link = (uint32_t)(s * 4 + mu);
            draw(seed, ST_DEMON, (uint64_t)sw, link, 3 * n, w0);
            draw(seed, ST_DEMON, (uint64_t)sw, link, 3 * n + 1, w1);
            draw(seed, ST_DEMON, (uint64_t)sw, link, 3 * n + 2, w2);
            double a0 = 0.9 + (1. - 0.9) * u01(w0[0], w0[1]);
            double a1 = -1. + 2. * u01(w0[2], w0[3]);
            double a2 = -1. + 2. * u01(w1[0], w1[1]);
            double a3 = -1. + 2. * u01(w1[2], w1[3]);
            double nrm = std::sqrt((1. - a0 * a0) / (a1 * a1 + a2 * a2 + a3 * a3));
            a1 *= nrm; a2 *= nrm; a3 *= nrm;
            Mat m = Mat::unit();
            switch (w2[0] % 3) {
              case 0: m(1, 1) = cplx(a0, a3); m(1, 2) = cplx(a2, a1); m(2, 1) = cplx(-a2, a1); m(2, 2) = cplx(a0, -a3); break;
              case 1: m(0, 0) = cplx(a0, a3); m(0, 2) = cplx(a2, a1); m(2, 0) = cplx(-a2, a1); m(2, 2) = cplx(a0, -a3); break;
              case 2: m(0, 0) = cplx(a0, a3); m(0, 1) = cplx(a2, a1); m(1, 0) = cplx(-a2, a1); m(1, 1) = cplx(a0, -a3); break;
            }
            Mat mv = m * v;
            double dS[4];
            bool ok = true;
            for (int q = 0; q < 4; ++q) {
              dS[q] = -retr(mv * X[q]) + retr(v * X[q]);
              ok = ok && (E[q] - dS[q] > -E_max) && (E[q] - dS[q] < E_max);
            }
            if (ok) {
              v = mv;
              for (int q = 0; q < 4; ++q) E[q] -= dS[q];
              ++acc;
            } else {
              ++rej;
            }
          }
        }
  store_field(u, U);
  acc_rej[0] = acc;
  acc_rej[1] = rej;
}

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
