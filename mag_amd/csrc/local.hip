// Constrained local Metropolis on the Wilson action: algGCtrnHeatBath::run
// (alg_gchb.h:290-387), FP64.
//
// Per iteration every link is updated by nHits Metropolis hits U -> U eps
// with eps = getRandomSU3BySubsu2(small) (alg_gchb.h:32-96), accepted with
// exp(S_old - S_new), S = -(beta/3) Re Tr(U sigma) and sigma the plaquette
// staple sum of the link. A constrained link (transverse coordinates
// multiples of mag) that is not the last of its coarse link is updated
// together with its successor: U1 eps, eps^+ U2 (metropolisKernelPProd,
// keeps U1 U2). After each iteration: reunitarize, plaquette and
// constrained plaquette.
//
// Two orders:
//  * reference: the reference's sequential order (mu, then sites with x0
//    fastest), executed by one lane; bit-for-bit comparable to the oracle;
//  * checkerboard: for each mu, four classes by (x_mu parity, parity of the
//    transverse coordinate sum). A link's staple holds mu-links only at
//    x +- nu (other transverse parity), a pair's second link sits at the
//    other x_mu parity and its staple at the other transverse parity, so
//    the updates of one class share no staple: each class is one launch,
//    one thread per member, and the result does not depend on thread order.
// Random numbers: Philox stream (seed, link, iteration); hit h of the update
// of link (x, mu) takes draws 3h, 3h+1, 3h+2 (a, b, c, d, select, die in
// [-1, 1), as CPS's LRG with SetInterval(1, -1)); a pair draws from its
// first link. Hit counts go to per-thread slots (no atomics), summed at the end.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "kernels.hpp"
#include "runtime.hpp"

using namespace mag;

namespace {

using D = double;
using LV = LinkView<double, 18>;

struct LocalArgs {
  Geom g;
  LV U;
  double fbeta;  // beta / 3
  double small;
  int nhits, mag;
  uint64_t seed, iter;
  double* cnt;   // [thread][4]: accept, reject, acceptCons, rejectCons
  int a, b;      // checkerboard class
};

MAG_D void shift_rt(const Geom& g, int* y, int k, int n) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i == k) {
      const int v = y[i] + n, L = g.L[i];
      y[i] = v < 0 ? v + L : (v >= L ? v - L : v);
    }
}
MAG_D int lin(const Geom& g, const int* y) { return y[0] + g.S[1] * y[1] + g.S[2] * y[2] + g.S[3] * y[3]; }

// MAG_LOCAL_READ12 = 1: staple links read as rows 0 and 1, row 2 rebuilt.
#ifndef MAG_LOCAL_READ12
#define MAG_LOCAL_READ12 1
#endif

// Plaquette staple sum sigma of link (x, mu): Re Tr(U_mu(x) sigma) is the sum
// of the six plaquettes through the link (get_staple_dagger, field-hmc.h:203-242).
MAG_D M3<D> staple(const LocalArgs& A, const int* x, int mu) {
  M3<D> s = zero3<D>();
#pragma unroll 1
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    int y[4] = {x[0], x[1], x[2], x[3]};
    shift_rt(A.g, y, mu, 1);
    const M3<D> a = A.U.template load_su3<D, MAG_LOCAL_READ12 != 0>(nu, lin(A.g, y));  // U_nu(x+mu)
    int z[4] = {x[0], x[1], x[2], x[3]};
    shift_rt(A.g, z, nu, 1);
    const M3<D> b = A.U.template load_su3<D, MAG_LOCAL_READ12 != 0>(mu, lin(A.g, z));  // U_mu(x+nu)
    const M3<D> c = A.U.template load_su3<D, MAG_LOCAL_READ12 != 0>(nu, lin(A.g, x));  // U_nu(x)
    mul_nd_acc(s, mul_nd(a, b), c);                          // U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
    shift_rt(A.g, y, nu, -1);                                // x+mu-nu
    int w[4] = {x[0], x[1], x[2], x[3]};
    shift_rt(A.g, w, nu, -1);                                // x-nu
    const M3<D> d = A.U.template load_su3<D, MAG_LOCAL_READ12 != 0>(nu, lin(A.g, y));  // U_nu(x+mu-nu)
    const M3<D> e = A.U.template load_su3<D, MAG_LOCAL_READ12 != 0>(mu, lin(A.g, w));  // U_mu(x-nu)
    const M3<D> f = A.U.template load_su3<D, MAG_LOCAL_READ12 != 0>(nu, lin(A.g, w));  // U_nu(x-nu)
    mul_acc(s, mul_dd(d, e), f);                             // U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
  }
  return s;
}

struct Draws {
  double u[6];
};
MAG_D Draws draws(const LocalArgs& A, uint32_t link, int hit) {
  Draws r;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const u32x4 w = rng_draw(A.seed, RNG_LOCAL, A.iter, link, (uint32_t)(3 * hit + k));
    r.u[2 * k] = -1. + 2. * rng_u01(w.x, w.y);
    r.u[2 * k + 1] = -1. + 2. * rng_u01(w.z, w.w);
  }
  return r;
}

// getRandomSU3BySubsu2 (alg_gchb.h:32-96)
MAG_D M3<D> subsu2(double small, const double* u) {
  double a = small * u[0];
  if (a > 0) a *= -1.;
  a += 1.;
  double b = u[1], c = u[2], d = u[3];
  const double norm = sqrt((1. - a * a) / (b * b + c * c + d * d));
  b *= norm;
  c *= norm;
  d *= norm;
  M3<D> m = zero3<D>();
  if (u[4] < -1. / 3.) {
    m(0, 0) = {a, b}; m(0, 1) = {c, d}; m(1, 0) = {-c, d}; m(1, 1) = {a, -b}; m(2, 2) = {1., 0.};
  } else if (u[4] > 1. / 3.) {
    m(0, 0) = {1., 0.}; m(1, 1) = {a, b}; m(1, 2) = {c, d}; m(2, 1) = {-c, d}; m(2, 2) = {a, -b};
  } else {
    m(0, 0) = {a, b}; m(0, 2) = {c, d}; m(1, 1) = {1., 0.}; m(2, 0) = {-c, d}; m(2, 2) = {a, -b};
  }
  return m;
}

MAG_D double die_of(const Draws& r) { return r.u[5] < 0. ? -r.u[5] : r.u[5]; }

// One link's update (alg_gchb.h:327-368): metropolisKernel, or
// metropolisKernelPProd with its successor for a constrained link below the last.
template <bool PAIR>
MAG_D void update_link(const LocalArgs& A, const int* x, int mu, double* c) {
  const int s = lin(A.g, x);
  const uint32_t link = (uint32_t)s * 4u + (uint32_t)mu;
  bool cons = A.mag > 0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i != mu) cons = cons && (x[i] % A.mag == 0);
  int xm = x[0];
#pragma unroll
  for (int i = 1; i < 4; ++i)
    if (i == mu) xm = x[i];
  if (PAIR && cons && xm % A.mag != A.mag - 1) {
    int y[4] = {x[0], x[1], x[2], x[3]};
    shift_rt(A.g, y, mu, 1);
    const int sy = lin(A.g, y);
    const M3<D> s1 = staple(A, x, mu), s2 = staple(A, y, mu);
    M3<D> U1 = A.U.template load_nt<D>(mu, s), U2 = A.U.template load_nt<D>(mu, sy);
    // the action of the kept links carries over from hit to hit
    double oldS = -A.fbeta * retrace_mul(U1, s1);
    oldS += -A.fbeta * retrace_mul(U2, s2);
#pragma unroll 1
    for (int h = 0; h < A.nhits; ++h) {
      const Draws r = draws(A, link, h);
      const M3<D> eps = subsu2(A.small, r.u);
      const M3<D> t1 = mul(U1, eps), t2 = mul_dn(eps, U2);
      double newS = -A.fbeta * retrace_mul(t1, s1);
      newS += -A.fbeta * retrace_mul(t2, s2);
      if (die_of(r) < exp(oldS - newS)) {
        U1 = t1;
        U2 = t2;
        oldS = newS;
        c[2] += 1.;
      } else {
        c[3] += 1.;
      }
    }
    A.U.store_nt(mu, s, U1);
    A.U.store_nt(mu, sy, U2);
  } else {
    const M3<D> sg = staple(A, x, mu);
    M3<D> U = A.U.template load_nt<D>(mu, s);
    double oldS = -A.fbeta * retrace_mul(U, sg);
#pragma unroll 1
    for (int h = 0; h < A.nhits; ++h) {
      const Draws r = draws(A, link, h);
      const M3<D> eps = subsu2(A.small, r.u);
      const M3<D> t1 = mul(U, eps);
      const double newS = -A.fbeta * retrace_mul(t1, sg);
      if (die_of(r) < exp(oldS - newS)) {
        U = t1;
        oldS = newS;
        c[0] += 1.;
      } else {
        c[1] += 1.;
      }
    }
    A.U.store_nt(mu, s, U);
  }
}

// reference order: one lane walks mu, then the sites in index order
__global__ __launch_bounds__(64) void k_local_reference(LocalArgs A) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double c[4] = {A.cnt[0], A.cnt[1], A.cnt[2], A.cnt[3]};
#pragma unroll 1
  for (int mu = 0; mu < 4; ++mu)
#pragma unroll 1
    for (int s = 0; s < A.g.V; ++s) {
      const Site p = site_of(A.g, s);
      update_link<true>(A, p.x, mu, c);
    }
#pragma unroll
  for (int k = 0; k < 4; ++k) A.cnt[k] = c[k];
}

// checkerboard class (MU, x_mu parity a, transverse parity b): thread t
// enumerates the V/4 members over a lattice halved along MU and along the
// first transverse direction T, which carries the parity fix-up.
// Without constrained pairs (mag = 0) the kernel is held to 2 waves per SIMD
// (20 bytes of spill; 7 % faster than 1 wave at 260 registers); the pair
// kernel stays at 1 wave (at 2 it spills 212 bytes and is 4 % slower).
template <int MU, bool PAIR>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(PAIR ? 1 : 2))) void k_local_class(LocalArgs A) {
  constexpr int T = (MU == 0) ? 1 : 0;
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= A.g.V / 4) return;
  int h[4];
  int r = t;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int Li = (i == MU || i == T) ? A.g.L[i] / 2 : A.g.L[i];
    h[i] = r % Li;
    r /= Li;
  }
  int x[4];
  int tr = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i == MU) x[i] = 2 * h[i] + A.a;
    else if (i == T) x[i] = 2 * h[i];
    else {
      x[i] = h[i];
      tr += h[i];
    }
  }
  x[T] += (A.b - tr) & 1;
  double c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) c[k] = A.cnt[(size_t)t * 4 + k];
  update_link<PAIR>(A, x, MU, c);
#pragma unroll
  for (int k = 0; k < 4; ++k) A.cnt[(size_t)t * 4 + k] = c[k];
}

int launch_local(mag_ctx c, void (*k)(LocalArgs), dim3 grid, dim3 block, const LocalArgs& a) {
  KTimer tm(c, "local_metropolis");
  hipLaunchKernelGGL(k, grid, block, 0, c->stream, a);
  CHECK_LAUNCH(c);
  return MAG_OK;
}

}  // namespace

// dst = w * (scalar slot 0), on the device
static int record_scalar(mag_ctx c, double* dst, double w) {
  LinCombArgs a{c->d_scalar, dst, 1, {}};
  a.w[0][0] = w;
  hipLaunchKernelGGL(k_lincomb<0>, dim3(1), dim3(64), 0, c->stream, a);
  CHECK_LAUNCH(c);
  return MAG_OK;
}

extern "C" {

int mag_local_metropolis(mag_field U, double beta, int mag, int num_iter, int nhits, double small, uint64_t seed,
                         int order, long acc_rej[4], double* plaq, double* cons_plaq) {
  if (!U || U->kind != MAG_LINK18 || U->prec != MAG_F64 || num_iter < 0 || nhits < 0 || mag < 0)
    return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  if (c->rec) return fail(c, MAG_ERR_UNSUPPORTED, "local Metropolis is not recordable");
  if (order != MAG_ORDER_REFERENCE && order != MAG_ORDER_CHECKERBOARD)
    return fail(c, MAG_ERR_ARG, "order must be MAG_ORDER_REFERENCE or MAG_ORDER_CHECKERBOARD");
  const Geom& g = U->g;
  if (order == MAG_ORDER_CHECKERBOARD)
    for (int i = 0; i < 4; ++i)
      if (g.L[i] % 2) return fail(c, MAG_ERR_GEOMETRY, "the checkerboard order needs even extents");
  if (mag > 0)
    for (int i = 0; i < 4; ++i)
      if (g.L[i] % mag) return fail(c, MAG_ERR_GEOMETRY, "extents must be multiples of mag");
  // the reference order is one lane walking every link: a check on small lattices
  if (order == MAG_ORDER_REFERENCE && g.V > 65536)
    return fail(c, MAG_ERR_UNSUPPORTED, "the reference order runs on one lane; use it up to 16^4 sites");
  const int nthr = order == MAG_ORDER_CHECKERBOARD ? g.V / 4 : 1;
  int r;
  // per-thread hit counts, in a buffer of the shape's workspace (the
  // reductions below use the context's partial-sum buffer)
  Workspace* ws = workspace(c, g, MAG_F64, 18);
  if ((r = ws->need_cnt())) return r;
  double* cnt = (double*)ws->Cnt;
  if ((r = fill_dev(c, cnt, 0., 4 * nthr))) return r;
  double* rec0 = nullptr;  // per iteration: plaquette, constrained plaquette
  if (plaq || cons_plaq) {
    if ((r = ws->need_rec(2 * (size_t)std::max(num_iter, 1)))) return r;
    rec0 = (double*)ws->Rec;
  }
  LocalArgs A{};
  A.g = g;
  A.U = LV{(double*)U->p, g.V};
  A.fbeta = beta / 3.;
  A.small = small;
  A.nhits = nhits;
  A.mag = mag;
  A.seed = seed;
  A.cnt = cnt;
  for (int it = 0; it < num_iter; ++it) {
    A.iter = (uint64_t)it;
    if (order == MAG_ORDER_REFERENCE) {
      if ((r = launch_local(c, k_local_reference, dim3(1), dim3(64), A))) return r;
    } else {
      const dim3 grid((g.V / 4 + BLOCK - 1) / BLOCK);
      // mag = 0 has no constrained pairs: that kernel holds one link and one
      // staple instead of two of each
      void (*kf[4])(LocalArgs) = {k_local_class<0, false>, k_local_class<1, false>, k_local_class<2, false>,
                                  k_local_class<3, false>};
      void (*kp[4])(LocalArgs) = {k_local_class<0, true>, k_local_class<1, true>, k_local_class<2, true>,
                                  k_local_class<3, true>};
      void (**ks)(LocalArgs) = mag > 0 ? kp : kf;
      for (int mu = 0; mu < 4; ++mu)
        for (A.a = 0; A.a < 2; ++A.a)
          for (A.b = 0; A.b < 2; ++A.b)
            if ((r = launch_local(c, ks[mu], grid, dim3(BLOCK), A))) return r;
    }
    // reunitarize, then record the plaquette and the constrained plaquette in
    // device memory (read once after the last iteration, not per iteration)
    if ((r = Ops<double, 18>::reunitarize(c, g, U->p, 0))) return r;
    if (plaq) {
      if ((r = Ops<double, 18>::smeared_action_sums(c, g, U->p, 0., 0, false, 0)) ||
          (r = record_scalar(c, rec0 + 2 * (size_t)it, 1. / (18. * (double)g.V))))
        return r;
    }
    if (cons_plaq && mag > 0) {
      double norm = 0.;
      if ((r = constrained_plaq_sum(c, g, MAG_F64, MAG_LINK18, U->p, mag, 0, &norm)) ||
          (r = record_scalar(c, rec0 + 2 * (size_t)it + 1, norm)))
        return r;
    }
  }
  if ((plaq || cons_plaq) && num_iter > 0) {
    std::vector<double> h((size_t)2 * num_iter);
    HIPCHK(c, hipMemcpyAsync(h.data(), rec0, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int it = 0; it < num_iter; ++it) {
      if (plaq) plaq[it] = h[2 * (size_t)it];
      if (cons_plaq) cons_plaq[it] = mag > 0 ? h[2 * (size_t)it + 1] : 0.;
    }
  }
  if (acc_rej) {
    // the per-thread counts summed on the device (whole numbers: exact in
    // any order), four scalars read back
    for (int k = 0; k < 4; ++k)
      if ((r = final_reduce(c, cnt + k, nthr, 4, 0, k))) return r;
    if ((r = fetch_scalars(c, 4))) return r;
    for (int k = 0; k < 4; ++k) acc_rej[k] = (long)c->h_scalar[k];
  }
  return MAG_OK;
}

}  // extern "C"
