// Coupled coarse/fine HMC of field-multi-md.h (namespace md::), FP64.
//
// The coarse lattice has half the fine extents; coarse site c sits at fine
// site x0 = 2c. For every coarse link (c, mu):
//   U2 = U_mu(x0) U_mu(x0+mu)                         get_U (field-multi.h:27-31)
//   Q  = -i rho TA(stp U2^+), stp = six 2x1 staples    get_Q (33-39), get_staple_rect
//   Gb = e^{iQ} U2                                     get_Gb (field-multi-md.h:771-783)
//   S_xi = -(1 + md_alpha) sum xi Re Tr(C^+ Gb)        get_xi_energy (30-72)
// The fine force (md::get_Fforce, 116-368) is the thin gauge force plus, on
// the links that enter some Gb, i TA(U mTemp) with mTemp collected from the
// Lambda-inserted loops of the reference's four stout_type classes. Each
// fine link of classes 1-4 belongs to exactly one (coarse site, slot) pair:
//   slot 0-3    class 1: x = 2c,             mu = slot
//   slot 4-7    class 3: x = 2c + mu,        mu = slot - 4
//   slot 8-19   class 2: x = 2c + nu,        (mu, nu) from slot - 8
//   slot 20-31  class 4: x = 2c + mu + nu,   (mu, nu) from slot - 20
// so the kernel's grid is (coarse sites, 32 slots): a wavefront works on one
// class and direction (no divergence) and every fine link is written by one
// thread (no atomics). Class-0 links (no Gb contains them) keep the gauge force.
// The coarse links' momentum (md::get_Cforce, 398-425) is the velocity of Gb
// along the fine momenta, -i alpha/(1+alpha) (dGb/dt) Gb^+; the loop
// derivatives (get_DQ, get_path_ordered_product_leftD) are carried along each
// loop as a (product, derivative) pair. Kernels here are not recordable
// (serving batches do not run this path).
#include <cmath>
#include <cstring>
#include <type_traits>

#include "kernels.hpp"
#include "runtime.hpp"

using namespace mag;

struct mag_multi_s {
  mag_ctx ctx = nullptr;
  int fdims[4] = {0, 0, 0, 0};
  mag_action act{};
  mag_hmc_params p{};
  mag_multi_params mp{};
  mag_field UF = nullptr, UC = nullptr, PF = nullptr, xi = nullptr;
  mag_field UFb = nullptr, UCb = nullptr;  // last accepted state
  long traj = 0;
};

namespace {

using D = double;
using LV = LinkView<double, 18>;
using AV = AlgView<double>;

#ifndef MAG_MULTI_WAVES
#define MAG_MULTI_WAVES 1
#endif
#define MAG_MULTI_OCC __attribute__((amdgpu_waves_per_eu(MAG_MULTI_WAVES)))
// occupancy floors of the fine-force kernels (heavy classes 1 and 3) and the coarse force
#ifndef MAG_FFORCE_WAVES
#define MAG_FFORCE_WAVES 1
#endif
#ifndef MAG_CFORCE_WAVES
#define MAG_CFORCE_WAVES 1
#endif
#define MAG_FFORCE_OCC __attribute__((amdgpu_waves_per_eu(MAG_FFORCE_WAVES)))
#define MAG_CFORCE_OCC __attribute__((amdgpu_waves_per_eu(MAG_CFORCE_WAVES)))

struct MultiArgs {
  Geom gf, gc;
  LV U;             // fine links
  LV C;             // coarse links
  const double* xi; // [mu][coarse site]
  AV P;             // fine momenta (coarse force)
  AV lam;           // coarse Lambda
  LV eq;            // coarse e^{iQ}
  AV F;             // fine force (+=) or coarse force (store)
  LV out;           // coarse Gb
  double rho, kappa, coeff;
  uint64_t seed, traj;
  double* part;
};

template <int K>
using IC = std::integral_constant<int, K>;

// A stored SU(3) link (fine or coarse link, e^{iQ}). MAG_MD_READ12 = 1
// reads rows 0 and 1 (6 of the 9 16-byte groups) and rebuilds row 2 as
// conj(r0 x r1): these kernels are bound by the vector-memory path, with the
// vector ALUs mostly idle.
#ifndef MAG_MD_READ12
#define MAG_MD_READ12 1
#endif
MAG_D M3<D> ld(const LV& v, int mu, int s) { return v.template load_su3<D, MAG_MD_READ12 != 0>(mu, s); }

// Shift of a fine coordinate along the compile-time direction K, wrapped
// into [0, L) (|n| <= 2 < L).
template <int K>
MAG_D void shift(const Geom& g, int* y, int n) {
  const int v = y[K] + n, L = g.L[K];
  y[K] = v < 0 ? v + L : (v >= L ? v - L : v);
}
// the same for a direction known only at run time (wave-uniform)
MAG_D void shift_rt(const Geom& g, int* y, int k, int n) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i == k) {
      const int v = y[i] + n, L = g.L[i];
      y[i] = v < 0 ? v + L : (v >= L ? v - L : v);
    }
}
MAG_D int lin(const Geom& g, const int* y) { return y[0] + g.S[1] * y[1] + g.S[2] * y[2] + g.S[3] * y[3]; }
// coarse index of an even fine coordinate
MAG_D int cidx(const Geom& gc, const int* y) {
  return (y[0] >> 1) + gc.S[1] * (y[1] >> 1) + gc.S[2] * (y[2] >> 1) + gc.S[3] * (y[3] >> 1);
}
MAG_D double xi_at(const MultiArgs& a, int mu, int sc) { return a.xi[(size_t)mu * a.gc.V + sc]; }

// f(IC<NU>) for the three directions NU != MU
template <int MU, class F>
MAG_D void for_nu(F&& f) {
  if constexpr (MU != 0) f(IC<0>{});
  if constexpr (MU != 1) f(IC<1>{});
  if constexpr (MU != 2) f(IC<2>{});
  if constexpr (MU != 3) f(IC<3>{});
}
// Directions are compile-time inside every kernel body: the block-uniform
// row of the work tile (ctile) selects a specialised body, so shifts and
// field-plane offsets fold to constants.
template <class F>
MAG_D void with_mu(int mu, F&& f) {
  switch (mu) {
    case 0: f(IC<0>{}); break;
    case 1: f(IC<1>{}); break;
    case 2: f(IC<2>{}); break;
    default: f(IC<3>{}); break;
  }
}
// k = 3 mu + (index of nu among the directions other than mu)
template <class F>
MAG_D void with_mu_nu(int k, F&& f) {
  switch (k) {
    case 0: f(IC<0>{}, IC<1>{}); break;
    case 1: f(IC<0>{}, IC<2>{}); break;
    case 2: f(IC<0>{}, IC<3>{}); break;
    case 3: f(IC<1>{}, IC<0>{}); break;
    case 4: f(IC<1>{}, IC<2>{}); break;
    case 5: f(IC<1>{}, IC<3>{}); break;
    case 6: f(IC<2>{}, IC<0>{}); break;
    case 7: f(IC<2>{}, IC<1>{}); break;
    case 8: f(IC<2>{}, IC<3>{}); break;
    case 9: f(IC<3>{}, IC<0>{}); break;
    case 10: f(IC<3>{}, IC<1>{}); break;
    default: f(IC<3>{}, IC<2>{}); break;
  }
}

// Ordered product along a fine path (get_path_ordered_product_insertion,
// field-matrix.h:177-200), starting at y.
struct Walk {
  LV u;
  Geom g;
  int y[4];
  M3<D> m;
  bool fresh = true;
  MAG_D Walk(const LV& u_, const Geom& g_, const int* x) : u(u_), g(g_) {
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] = x[i];
  }
  MAG_D void right(const M3<D>& a) {
    if (fresh) m = a;
    else m = mul(m, a);
    fresh = false;
  }
  template <int K>
  MAG_D void fwd() {
    right(ld(u, K, lin(g, y)));
    shift<K>(g, y, 1);
  }
  template <int K>
  MAG_D void bwd() {
    shift<K>(g, y, -1);
    const M3<D> a = ld(u, K, lin(g, y));
    if (fresh) m = dag(a);
    else m = mul_nd(m, a);
    fresh = false;
  }
  MAG_D void ins(const M3<D>& a) { right(a); }
};

// i-weighted accumulation: t += (i w) m
MAG_D void acc_i(M3<D>& t, double w, const M3<D>& m) { caxpy(t, cpx<D>{0., w}, m); }

// Q of coarse link (c, MU) (as algebra components) and U2; fine x0 = 2c.
template <int MU>
MAG_D void coarse_Q(const MultiArgs& a, const int* x0, D q[8], M3<D>& U2) {
  int y[4] = {x0[0], x0[1], x0[2], x0[3]};
  const M3<D> U1 = ld(a.U, MU, lin(a.gf, y));
  shift<MU>(a.gf, y, 1);
  U2 = mul(U1, ld(a.U, MU, lin(a.gf, y)));
  M3<D> stp = zero3<D>();
  for_nu<MU>([&](auto nc) {
    constexpr int NU = decltype(nc)::value;
    {  // nu, mu, mu, -nu
      Walk w(a.U, a.gf, x0);
      w.fwd<NU>(); w.fwd<MU>(); w.fwd<MU>(); w.bwd<NU>();
      add_to(stp, w.m);
    }
    {  // -nu, mu, mu, nu
      Walk w(a.U, a.gf, x0);
      w.bwd<NU>(); w.fwd<MU>(); w.fwd<MU>(); w.fwd<NU>();
      add_to(stp, w.m);
    }
  });
  force_alg_from(mul_nd(stp, U2), q);  // -i TA(stp U2^+)
#pragma unroll
  for (int k = 0; k < 8; ++k) q[k] *= a.rho;
}

// Work tile of a 1D grid of nb * ny blocks: nb blocks of 256 coarse sites
// times ny rows (directions, or (mu, nu) pairs). Blocks go to the 8 XCDs
// round-robin; when nb splits into 8, XCD k owns the coarse blocks
// [k nb/8, (k+1) nb/8) (a slab of x3, which is slowest in the site order)
// and walks them with the row fastest, so the rows of one site block run
// back to back on one XCD and the fine links they share come into its L2
// once; the slab's neighbours along x3 are on at most two other XCDs.
// Otherwise row-major order.
struct CTile {
  int row, blk;
};
MAG_D CTile ctile(const MultiArgs& a) {
  const int nb = (a.gc.V + BLOCK - 1) / BLOCK, ny = gridDim.x / nb, b = blockIdx.x;
  if (nb % 8 == 0) {
    const int j = b >> 3;
    return {j % ny, (b & 7) * (nb >> 3) + j / ny};
  }
  return {b / nb, b % nb};
}

MAG_D void coarse_site(const MultiArgs& a, int sc, int* x0) {
  const Site pc = site_of(a.gc, sc);
#pragma unroll
  for (int i = 0; i < 4; ++i) x0[i] = 2 * pc.x[i];
}

// Lambda (compute_Lambda with Sigma' = C^+, field-multi-md.h:164-176) and e^{iQ}.
__global__ __launch_bounds__(BLOCK) MAG_MULTI_OCC void k_multi_prep(MultiArgs a) {
  const CTile tl = ctile(a);
  const int sc = tl.blk * BLOCK + threadIdx.x;
  if (sc >= a.gc.V) return;
  int x0[4];
  coarse_site(a, sc, x0);
  with_mu(tl.row, [&](auto mc) {
    constexpr int MU = decltype(mc)::value;
    D q[8];
    M3<D> U2;
    coarse_Q<MU>(a, x0, q, U2);
    const M3<D> X = mul_nd(U2, ld(a.C, MU, sc));  // U2 Sigma'
    D lam[8];
    M3<D> eQ;
    ch_lambda(alg_to_mat(q), X, lam, eQ);
    a.lam.store(MU, sc, lam);
    a.eq.store(MU, sc, eQ);
  });
}

// ---------------------------------------------------------------- fine force
// The class bodies follow field-multi-md.h:196-357 path by path; y is the
// insertion position of each path, c(y) its coarse site.

// class 1 (196-252): x = 2c
template <int MU>
MAG_D M3<D> fclass1(const MultiArgs& a, const int* x, int sc) {
  const D r = a.rho;
  int s[4] = {x[0], x[1], x[2], x[3]};
  shift<MU>(a.gf, s, 1);
  const D xm = xi_at(a, MU, sc);
  const M3<D> Lm = a.lam.template load_mat<D>(MU, sc);
  // direct term U(x+mu) C^+ e^{iQ} xi
  M3<D> t = scale(mul(mul_nd(ld(a.U, MU, lin(a.gf, s)), ld(a.C, MU, sc)),
                      ld(a.eq, MU, sc)),
                  xm);
  for_nu<MU>([&](auto nc) {
    constexpr int NU = decltype(nc)::value;
    {
      Walk w(a.U, a.gf, s);
      w.fwd<MU>(); w.fwd<NU>(); w.bwd<MU>(); w.bwd<MU>(); w.bwd<NU>(); w.ins(Lm);
      acc_i(t, r * xm, w.m);
    }
    {
      Walk w(a.U, a.gf, s);
      w.fwd<MU>(); w.bwd<NU>(); w.bwd<MU>(); w.bwd<MU>(); w.fwd<NU>(); w.ins(Lm);
      acc_i(t, r * xm, w.m);
    }
    {
      Walk w(a.U, a.gf, s);
      w.fwd<NU>(); w.fwd<NU>(); w.bwd<MU>(); w.bwd<NU>(); w.bwd<NU>(); w.ins(a.lam.template load_mat<D>(NU, sc));
      acc_i(t, -r * xi_at(a, NU, sc), w.m);
    }
    {
      int y[4] = {x[0], x[1], x[2], x[3]};
      shift<NU>(a.gf, y, -2);
      const int cy = cidx(a.gc, y);
      Walk w(a.U, a.gf, s);
      w.bwd<NU>(); w.bwd<NU>(); w.bwd<MU>(); w.ins(a.lam.template load_mat<D>(NU, cy)); w.fwd<NU>(); w.fwd<NU>();
      acc_i(t, r * xi_at(a, NU, cy), w.m);
    }
  });
  return t;
}

// class 2 (254-278): x = 2c + nu
template <int MU, int NU>
MAG_D M3<D> fclass2(const MultiArgs& a, const int* x, int sc) {
  const D r = a.rho;
  int s[4] = {x[0], x[1], x[2], x[3]};
  shift<MU>(a.gf, s, 1);
  M3<D> t = zero3<D>();
  {
    Walk w(a.U, a.gf, s);
    w.fwd<MU>(); w.bwd<NU>(); w.bwd<MU>(); w.bwd<MU>(); w.ins(a.lam.template load_mat<D>(MU, sc)); w.fwd<NU>();
    acc_i(t, -r * xi_at(a, MU, sc), w.m);
  }
  {
    int y[4] = {x[0], x[1], x[2], x[3]};
    shift<NU>(a.gf, y, 1);
    const int cy = cidx(a.gc, y);
    Walk w(a.U, a.gf, s);
    w.fwd<MU>(); w.fwd<NU>(); w.bwd<MU>(); w.bwd<MU>(); w.ins(a.lam.template load_mat<D>(MU, cy)); w.bwd<NU>();
    acc_i(t, -r * xi_at(a, MU, cy), w.m);
  }
  return t;
}

// class 3 (279-332): x = 2c + mu
template <int MU>
MAG_D M3<D> fclass3(const MultiArgs& a, const int* x, int sc) {
  const D r = a.rho;
  int s[4] = {x[0], x[1], x[2], x[3]};
  shift<MU>(a.gf, s, 1);
  int ym[4] = {x[0], x[1], x[2], x[3]};
  shift<MU>(a.gf, ym, -1);  // = 2c
  const D xm = xi_at(a, MU, sc);
  const M3<D> Lm = a.lam.template load_mat<D>(MU, sc);
  const M3<D> Uym = ld(a.U, MU, lin(a.gf, ym));
  // direct term C^+ e^{iQ} U(x-mu) xi
  M3<D> t = scale(mul(mul_dn(ld(a.C, MU, sc), ld(a.eq, MU, sc)), Uym), xm);
  int yp[4] = {x[0], x[1], x[2], x[3]};
  shift<MU>(a.gf, yp, 1);  // = 2(c + mu)
  const int cp = cidx(a.gc, yp);
  for_nu<MU>([&](auto nc) {
    constexpr int NU = decltype(nc)::value;
    {
      Walk w(a.U, a.gf, s);
      w.fwd<NU>(); w.bwd<MU>(); w.bwd<MU>(); w.bwd<NU>(); w.ins(Lm); w.right(Uym);
      acc_i(t, r * xm, w.m);
    }
    {
      Walk w(a.U, a.gf, s);
      w.bwd<NU>(); w.bwd<MU>(); w.bwd<MU>(); w.fwd<NU>(); w.ins(Lm); w.right(Uym);
      acc_i(t, r * xm, w.m);
    }
    {
      Walk w(a.U, a.gf, s);
      w.ins(a.lam.template load_mat<D>(NU, cp)); w.fwd<NU>(); w.fwd<NU>(); w.bwd<MU>(); w.bwd<NU>(); w.bwd<NU>();
      acc_i(t, r * xi_at(a, NU, cp), w.m);
    }
    {
      int y[4] = {yp[0], yp[1], yp[2], yp[3]};
      shift<NU>(a.gf, y, -2);
      const int cy = cidx(a.gc, y);
      Walk w(a.U, a.gf, s);
      w.bwd<NU>(); w.bwd<NU>(); w.ins(a.lam.template load_mat<D>(NU, cy)); w.bwd<MU>(); w.fwd<NU>(); w.fwd<NU>();
      acc_i(t, -r * xi_at(a, NU, cy), w.m);
    }
  });
  return t;
}

// class 4 (333-357): x = 2c + mu + nu
template <int MU, int NU>
MAG_D M3<D> fclass4(const MultiArgs& a, const int* x, int sc) {
  const D r = a.rho;
  int s[4] = {x[0], x[1], x[2], x[3]};
  shift<MU>(a.gf, s, 1);
  M3<D> t = zero3<D>();
  {
    Walk w(a.U, a.gf, s);
    w.bwd<NU>(); w.bwd<MU>(); w.bwd<MU>(); w.ins(a.lam.template load_mat<D>(MU, sc)); w.fwd<NU>(); w.fwd<MU>();
    acc_i(t, -r * xi_at(a, MU, sc), w.m);
  }
  {
    int y[4] = {x[0], x[1], x[2], x[3]};
    shift<NU>(a.gf, y, 1);
    shift<MU>(a.gf, y, -1);
    const int cy = cidx(a.gc, y);
    Walk w(a.U, a.gf, s);
    w.fwd<NU>(); w.bwd<MU>(); w.bwd<MU>(); w.ins(a.lam.template load_mat<D>(MU, cy)); w.bwd<NU>(); w.fwd<MU>();
    acc_i(t, -r * xi_at(a, MU, cy), w.m);
  }
  return t;
}

// fx += i TA(U(x) mTemp) (363-365); i TA(W) has components -Im Tr(T_a W)
template <int MU>
MAG_D void add_stout_term(const MultiArgs& a, const int* x, const M3<D>& t) {
  const int s = lin(a.gf, x);
  D w[8], f[8];
  force_alg_from(mul(ld(a.U, MU, s), t), w);
  a.F.load(MU, s, f);
#pragma unroll
  for (int k = 0; k < 8; ++k) f[k] -= w[k];
  a.F.store(MU, s, f);
}

// One launch of 32 rows: rows 0-3 class 1 and 4-7 class 3 (direction row & 3),
// rows 8-19 class 2 and 20-31 class 4 ((mu, nu) pair (row - 8) % 12). ctile
// puts the 32 rows of a coarse block back to back on one XCD, so all classes
// of a coarse site read the fine links they share from its L2 (one launch per
// class took 1.4x as long, one per pair of classes 1.05x, though the light
// classes would run 2 waves per SIMD on their own).
__global__ __launch_bounds__(BLOCK) MAG_FFORCE_OCC void k_multi_fforce(MultiArgs a) {
  const CTile tl = ctile(a);
  const int sc = tl.blk * BLOCK + threadIdx.x;
  if (sc >= a.gc.V) return;
  int x[4];
  coarse_site(a, sc, x);
  if (tl.row < 8) {
    with_mu(tl.row & 3, [&](auto mc) {
      constexpr int MU = decltype(mc)::value;
      if (tl.row < 4) {
        add_stout_term<MU>(a, x, fclass1<MU>(a, x, sc));
      } else {
        shift<MU>(a.gf, x, 1);
        add_stout_term<MU>(a, x, fclass3<MU>(a, x, sc));
      }
    });
  } else {
    const int r = tl.row - 8;
    with_mu_nu(r < 12 ? r : r - 12, [&](auto mc, auto nc) {
      constexpr int MU = decltype(mc)::value, NU = decltype(nc)::value;
      shift<NU>(a.gf, x, 1);
      if (r < 12) {
        add_stout_term<MU>(a, x, fclass2<MU, NU>(a, x, sc));
      } else {
        shift<MU>(a.gf, x, 1);
        add_stout_term<MU>(a, x, fclass4<MU, NU>(a, x, sc));
      }
    });
  }
}

// ---------------------------------------------------------------- coarse force
// Loop product and its derivative along U' = i P U on every link
// (get_path_ordered_product_leftD, field-matrix.h:138-175), carried as the
// pair (m, d) = (product, derivative) of the path so far:
//   forward link X = U:     d <- (d + i m P) X,        m <- m X
//   backward link X = U^+:  d <- d X - i (m X) P,      m <- m X
// (three 3x3 products per link).
struct DWalk {
  LV u;
  AV p;
  Geom g;
  int y[4];
  M3<D> m, d;
  bool fresh = true;
  MAG_D DWalk(const MultiArgs& a, const int* x) : u(a.U), p(a.P), g(a.gf) {
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] = x[i];
  }
  MAG_D void fwd(int K) {
    const int s = lin(g, y);
    const M3<D> X = ld(u, K, s);
    const M3<D> P = p.template load_mat<D>(K, s);
    if (fresh) {
      m = X;
      d = cscale(mul_hn(P, X), cpx<D>{0., 1.});
      fresh = false;
    } else {
      M3<D> mP = zero3<D>();
      mul_acc_nh(mP, m, P);
#pragma unroll
      for (int k = 0; k < 9; ++k) d.e[k] = {d.e[k].r - mP.e[k].i, d.e[k].i + mP.e[k].r};
      d = mul(d, X);
      m = mul(m, X);
    }
    shift_rt(g, y, K, 1);
  }
  MAG_D void bwd(int K) {
    shift_rt(g, y, K, -1);
    const int s = lin(g, y);
    const M3<D> Uy = ld(u, K, s);
    const M3<D> P = p.template load_mat<D>(K, s);
    if (fresh) {
      m = dag(Uy);
      d = cscale(mul_dh(Uy, P), cpx<D>{0., -1.});
      fresh = false;
    } else {
      const M3<D> mX = mul_nd(m, Uy);
      M3<D> mXP = zero3<D>();
      mul_acc_nh(mXP, mX, P);
      d = mul_nd(d, Uy);
#pragma unroll
      for (int k = 0; k < 9; ++k) d.e[k] = {d.e[k].r + mXP.e[k].i, d.e[k].i - mXP.e[k].r};
      m = mX;
    }
  }
};

template <int MU>
MAG_D void cforce_link(const MultiArgs& a, const int* x0, int sc) {
  // -i TA of Omega = stp U2^+ and of its derivative over the 6 rectangle
  // loops (get_DQ). Every loop is a 2x1 staple from x0 to x0 + 2 mu closed by
  // the same two links back, and the (product, derivative) step is linear in
  // the pair, so the staples and their derivatives are summed first and the
  // closing links applied once. The nu loop stays rolled (one nu's paths
  // live at a time).
  M3<D> S = zero3<D>(), dS = zero3<D>();
#pragma unroll 1
  for (int j = 0; j < 3; ++j) {
    const int nu = j < MU ? j : j + 1;
    {
      DWalk w(a, x0);
      w.fwd(nu); w.fwd(MU); w.fwd(MU); w.bwd(nu);
      add_to(S, w.m);
      add_to(dS, w.d);
    }
    {
      DWalk w(a, x0);
      w.bwd(nu); w.fwd(MU); w.fwd(MU); w.fwd(nu);
      add_to(S, w.m);
      add_to(dS, w.d);
    }
  }
  D q[8], dq[8];
  {
    DWalk w(a, x0);
    shift<MU>(a.gf, w.y, 2);
    w.m = S;
    w.d = dS;
    w.fresh = false;
    w.bwd(MU);
    w.bwd(MU);
    force_alg_from(w.m, q);
    force_alg_from(w.d, dq);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    q[k] *= a.rho;
    dq[k] *= a.rho;
  }
  const M3<D> Q = alg_to_mat(q), DQ = alg_to_mat(dq);
  // U2 and its derivative
  DWalk w(a, x0);
  w.fwd(MU);
  w.fwd(MU);
  const M3<D> U2 = w.m, DU2 = w.d;
  // DexpiQ (field-multi.h:90-178); zero when c1 == 0, as in the reference
  M3<D> E, dE;
  dexpiq(Q, DQ, E, dE);
  // (DexpiQ U2 + e^{iQ} DU2) (e^{iQ} U2)^+, times -i alpha/(1+alpha)
  M3<D> V = mul(dE, U2);
  mul_acc(V, E, DU2);
  const M3<D> G = mul(E, U2);
  D h[8];
  force_alg_from(mul_nd(V, G), h);  // Re Tr(T_a (-i V G^+))
#pragma unroll
  for (int k = 0; k < 8; ++k) h[k] *= a.kappa;
  a.F.store(MU, sc, h);
}


// ---------------------------------------------------------------- Gb, S_xi, heat bath
template <int MU>
MAG_D M3<D> coarse_Gb(const MultiArgs& a, const int* x0) {
  D q[8];
  M3<D> U2;
  coarse_Q<MU>(a, x0, q, U2);
  return mul(expiQ(alg_to_mat(q)), U2);
}

__global__ __launch_bounds__(BLOCK) MAG_MULTI_OCC void k_multi_gb(MultiArgs a) {
  const CTile tl = ctile(a);
  const int sc = tl.blk * BLOCK + threadIdx.x;
  if (sc >= a.gc.V) return;
  int x0[4];
  coarse_site(a, sc, x0);
  with_mu(tl.row, [&](auto mc) {
    constexpr int MU = decltype(mc)::value;
    a.out.store(MU, sc, coarse_Gb<MU>(a, x0));
  });
}

// get_Q (field-multi.h:33-39) per coarse link, stored as its 8 generator components
__global__ __launch_bounds__(BLOCK) MAG_MULTI_OCC void k_multi_getq(MultiArgs a) {
  const CTile tl = ctile(a);
  const int sc = tl.blk * BLOCK + threadIdx.x;
  if (sc >= a.gc.V) return;
  int x0[4];
  coarse_site(a, sc, x0);
  with_mu(tl.row, [&](auto mc) {
    constexpr int MU = decltype(mc)::value;
    D q[8];
    M3<D> U2;
    coarse_Q<MU>(a, x0, q, U2);
    a.F.store(MU, sc, q);
  });
}

// per-link terms xi Re Tr(C^+ Gb), one direction per grid row (summed by the
// final reduction over the partial sums of all rows)
__global__ __launch_bounds__(BLOCK) MAG_MULTI_OCC void k_multi_xi(MultiArgs a) {
  const CTile tl = ctile(a);
  const int sc = tl.blk * BLOCK + threadIdx.x;
  double acc = 0.;
  if (sc < a.gc.V) {
    int x0[4];
    coarse_site(a, sc, x0);
    with_mu(tl.row, [&](auto mc) {
      constexpr int MU = decltype(mc)::value;
      acc = retrace_mul_nd(coarse_Gb<MU>(a, x0), ld(a.C, MU, sc)) * xi_at(a, MU, sc);
    });
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) a.part[tl.row * ((a.gc.V + BLOCK - 1) / BLOCK) + tl.blk] = t;
}

// simplest_metropolis with random_su3_from_su2(0.9) (field-multi-md.h:748-769,
// field-matrix.h:36-65); Philox heat-bath stream keyed by (coarse link, traj).
template <int MU>
MAG_D void heatbath_link(const MultiArgs& a, const int* x0, int sc) {
  const M3<D> env = coarse_Gb<MU>(a, x0);
  const uint32_t link = (uint32_t)sc * 4u + (uint32_t)MU;
  const u32x4 w0 = rng_draw(a.seed, RNG_HEATBATH, a.traj, link, 0);
  const u32x4 w1 = rng_draw(a.seed, RNG_HEATBATH, a.traj, link, 1);
  const u32x4 w2 = rng_draw(a.seed, RNG_HEATBATH, a.traj, link, 2);
  const double a0 = 0.9 + (1. - 0.9) * rng_u01(w0.x, w0.y);
  double a1 = -1. + 2. * rng_u01(w0.z, w0.w);
  double a2 = -1. + 2. * rng_u01(w1.x, w1.y);
  double a3 = -1. + 2. * rng_u01(w1.z, w1.w);
  const double nrm = sqrt((1. - a0 * a0) / (a1 * a1 + a2 * a2 + a3 * a3));
  a1 *= nrm;
  a2 *= nrm;
  a3 *= nrm;
  M3<D> m = unit3<D>();
  switch (w2.x % 3) {
    case 0: m(1, 1) = {a0, a3}; m(1, 2) = {a2, a1}; m(2, 1) = {-a2, a1}; m(2, 2) = {a0, -a3}; break;
    case 1: m(0, 0) = {a0, a3}; m(0, 2) = {a2, a1}; m(2, 0) = {-a2, a1}; m(2, 2) = {a0, -a3}; break;
    default: m(0, 0) = {a0, a3}; m(0, 1) = {a2, a1}; m(1, 0) = {-a2, a1}; m(1, 1) = {a0, -a3}; break;
  }
  const M3<D> var = ld(a.C, MU, sc);
  const M3<D> nv = mul(var, m);
  const double diff = a.coeff * retrace(sub(mul_nd(nv, env), mul_nd(var, env)));
  if (rng_u01(w2.z, w2.w) < exp(diff)) a.C.store(MU, sc, nv);
}

__global__ __launch_bounds__(BLOCK) MAG_MULTI_OCC void k_multi_heatbath(MultiArgs a) {
  const CTile tl = ctile(a);
  const int sc = tl.blk * BLOCK + threadIdx.x;
  if (sc >= a.gc.V) return;
  int x0[4];
  coarse_site(a, sc, x0);
  with_mu(tl.row, [&](auto mc) { heatbath_link<decltype(mc)::value>(a, x0, sc); });
}

// ---------------------------------------------------------------- host side
bool f64link(mag_field f) { return f && f->kind == MAG_LINK18 && f->prec == MAG_F64; }
bool f64alg(mag_field f) { return f && f->kind == MAG_ALG8 && f->prec == MAG_F64; }
bool f64real(mag_field f) { return f && f->kind == MAG_REAL1 && f->prec == MAG_F64; }
bool dims_eq(mag_field f, const int* d) {
  for (int i = 0; i < 4; ++i)
    if (f->dims[i] != d[i]) return false;
  return true;
}

struct Multi {
  mag_ctx c;
  Geom gf, gc;
  Workspace *wf, *wc;
  mag_multi_params mp;
  int mag;
};

// Validates the fine links and derives the coarse geometry.
int make_multi(mag_field UF, const mag_multi_params* mp, int mag, Multi* m) {
  if (!f64link(UF) || !mp) return MAG_ERR_ARG;
  mag_ctx c = UF->ctx;
  int dc[4];
  for (int i = 0; i < 4; ++i) {
    if (UF->dims[i] % 2) return fail(c, MAG_ERR_GEOMETRY, "fine extents of the coarse/fine HMC must be even");
    dc[i] = UF->dims[i] / 2;
  }
  m->c = c;
  m->gf = UF->g;
  m->gc = make_geom(dc);
  m->wf = workspace(c, m->gf, MAG_F64, 18);
  m->wc = workspace(c, m->gc, MAG_F64, 18);
  m->mp = *mp;
  m->mag = mag;
  return MAG_OK;
}

int check_coarse(const Multi& m, mag_field f, bool (*kind)(mag_field)) {
  if (!kind(f) || f->ctx != m.c) return fail(m.c, MAG_ERR_ARG, "coarse field of the wrong kind or context");
  for (int i = 0; i < 4; ++i)
    if (f->dims[i] != m.gc.L[i]) return fail(m.c, MAG_ERR_GEOMETRY, "coarse extents must be half the fine extents");
  return MAG_OK;
}
int check_fine(const Multi& m, mag_field f, bool (*kind)(mag_field)) {
  if (!kind(f) || f->ctx != m.c) return fail(m.c, MAG_ERR_ARG, "fine field of the wrong kind or context");
  for (int i = 0; i < 4; ++i)
    if (f->dims[i] != m.gf.L[i]) return fail(m.c, MAG_ERR_GEOMETRY, "fine fields differ in extents");
  return MAG_OK;
}

MultiArgs args(const Multi& m, void* U, void* C, const void* xi) {
  MultiArgs a{};
  a.gf = m.gf;
  a.gc = m.gc;
  a.U = LV{(double*)U, m.gf.V};
  a.C = LV{(double*)C, m.gc.V};
  a.xi = (const double*)xi;
  a.rho = m.mp.rho;
  a.kappa = m.mp.md_alpha / (1. + m.mp.md_alpha);
  a.coeff = (1. + m.mp.md_alpha) * m.mp.xi0;
  return a;
}

int mlaunch(const Multi& m, const char* name, void (*k)(MultiArgs), dim3 grid, const MultiArgs& a) {
  if (m.c->rec) return fail(m.c, MAG_ERR_UNSUPPORTED, "the coarse/fine HMC kernels are not recordable");
  KTimer t(m.c, name);
  hipLaunchKernelGGL(k, grid, dim3(BLOCK), 0, m.c->stream, a);
  CHECK_LAUNCH(m.c);
  return MAG_OK;
}

// 1D grid of (coarse site block, row) tiles; see ctile()
dim3 cgrid(const Multi& m, int y) { return dim3((unsigned)(((m.gc.V + BLOCK - 1) / BLOCK) * y), 1, 1); }

// md::get_Fforce into F (fine alg8)
int fforce(const Multi& m, void* F, void* U, void* C, const void* xi, const mag_action& act) {
  int r;
  if ((r = m.wc->need_multi())) return r;
  mag_action thin = act;
  thin.mag = 0;
  thin.n_stout = 0;
  // the thin gauge force, without link classes (130-159)
  if ((r = Ops<double, 18>::force(m.c, m.gf, U, thin, OUT_F, F, nullptr, nullptr, 0., m.wf))) return r;
  MultiArgs a = args(m, U, C, xi);
  a.lam = AV{(double*)m.wc->Lam, m.gc.V};
  a.eq = LV{(double*)m.wc->Eq, m.gc.V};
  a.F = AV{(double*)F, m.gf.V};
  if ((r = mlaunch(m, "multi_prep", k_multi_prep, cgrid(m, 4), a))) return r;
  return mlaunch(m, "multi_fforce", k_multi_fforce, cgrid(m, 32), a);
}

// The four directions are the rows of one launch, so ctile puts a coarse
// block's four directions back to back on one XCD and the fine links and
// momenta they share are fetched into its L2 once (one launch per direction
// took 30 % longer).
__global__ __launch_bounds__(BLOCK) MAG_CFORCE_OCC void k_multi_cforce(MultiArgs a) {
  const CTile tl = ctile(a);
  const int sc = tl.blk * BLOCK + threadIdx.x;
  if (sc >= a.gc.V) return;
  int x0[4];
  coarse_site(a, sc, x0);
  with_mu(tl.row, [&](auto mc) { cforce_link<decltype(mc)::value>(a, x0, sc); });
}

// md::get_Cforce into CF (coarse alg8)
int cforce(const Multi& m, void* CF, void* U, void* P) {
  MultiArgs a = args(m, U, nullptr, nullptr);
  a.P = AV{(double*)P, m.gf.V};
  a.F = AV{(double*)CF, m.gc.V};
  return mlaunch(m, "multi_cforce", k_multi_cforce, cgrid(m, 4), a);
}


// sum of xi Re Tr(C^+ Gb) into scalar slot
int xi_sum(const Multi& m, void* U, void* C, const void* xi, int slot) {
  const int nb = (m.gc.V + BLOCK - 1) / BLOCK;
  int r;
  if ((r = ensure_part(m.c, 4 * (size_t)nb))) return r;
  MultiArgs a = args(m, U, C, xi);
  a.part = m.c->d_part;
  if ((r = mlaunch(m, "multi_xi", k_multi_xi, cgrid(m, 4), a))) return r;
  return final_reduce(m.c, m.c->d_part, 4 * nb, 1, 0, slot);
}

// Link updates C <- e^{i eps Cm} Cin and U <- e^{i eps P} Uin (Uin = U unless
// given: an update out of a copy source replaces the copy). The links are
// read through ld()'s rule (rows 0 and 1 with MAG_MD_READ12).
int md_update(const Multi& m, const Geom& g, const void* Uin, void* Uout, void* P, double eps) {
  if (m.c->rec) return fail(m.c, MAG_ERR_UNSUPPORTED, "the coarse/fine HMC kernels are not recordable");
  const UpdateArgs<double, 18> a{g, LV{(double*)Uin, g.V}, LV{(double*)Uout, g.V}, AV{(double*)P, g.V}, eps, m.mag};
  KTimer t(m.c, "link_update");
  hipLaunchKernelGGL((k_link_update<double, 18, double, MAG_MD_READ12 != 0>), grid_links(g), dim3(BLOCK), 0,
                     m.c->stream, a);
  CHECK_LAUNCH(m.c);
  return MAG_OK;
}
int coarse_update(const Multi& m, void* C, void* Cm, double eps, const void* Cin = nullptr) {
  return md_update(m, m.gc, Cin ? Cin : C, C, Cm, eps);
}
int fine_update(const Multi& m, void* U, void* P, double eps, const void* Uin = nullptr) {
  return md_update(m, m.gf, Uin ? Uin : U, U, P, eps);
}

// force_gradient_integrator_zeta_T (field-multi-md.h:441-475). With U0 / C0
// the fine / coarse fields start as copies of U0 / C0 (the first update of
// each reads them).
int zeta_T(const Multi& m, void* U, void* P, void* C, void* Cm, double dt, const void* U0 = nullptr,
           const void* C0 = nullptr) {
  int r;
  if ((r = fine_update(m, U, P, dt / 6., U0))) return r;
  if ((r = cforce(m, Cm, U, P)) || (r = coarse_update(m, C, Cm, 0.5 * dt, C0))) return r;
  if ((r = fine_update(m, U, P, 2. * dt / 3.))) return r;
  if ((r = cforce(m, Cm, U, P)) || (r = coarse_update(m, C, Cm, 0.5 * dt))) return r;
  return fine_update(m, U, P, dt / 6.);
}

int reunit_both(const Multi& m, void* U, void* C, int slot) {
  int r = Ops<double, 18>::reunitarize(m.c, m.gf, U, slot);
  if (r) return r;
  return Ops<double, 18>::reunitarize(m.c, m.gc, C, slot + 1);
}

// force_gradient_integrator_zeta_H (477-597); reunitarize deviations to slots (dslot, dslot+1)
int zeta_H(const Multi& m, void* U, void* P, void* C, const void* xi, const mag_action& act, int steps, double dt,
           int dslot) {
  int r;
  if ((r = m.wf->need_md()) || (r = m.wc->need_md())) return r;
  void *Ff = m.wf->F, *FgA = m.wf->Uaux, *CgA = m.wc->Uaux, *Cm = m.wc->F;
  const double alpha = (3. - std::sqrt(3.)) * dt / 6.;
  const double beta = dt / std::sqrt(3.);
  const double gamma = (2. - std::sqrt(3.)) * dt * dt / 12.;
  if ((r = zeta_T(m, U, P, C, Cm, alpha))) return r;
  for (int i = 0; i < steps; ++i) {
    for (int h = 0; h < 2; ++h) {
      if ((r = fforce(m, Ff, U, C, xi, act))) return r;
      if ((r = zeta_T(m, FgA, Ff, CgA, Cm, gamma, U, C))) return r;  // FgA, CgA from U, C
      if ((r = fforce(m, Ff, FgA, CgA, xi, act))) return r;
      if ((r = mom_update(m.c, m.gf, MAG_F64, P, Ff, 0.5 * dt))) return r;
      if (h == 0 && (r = zeta_T(m, U, P, C, Cm, beta))) return r;
    }
    if ((r = zeta_T(m, U, P, C, Cm, i < steps - 1 ? 2. * alpha : alpha))) return r;
  }
  return reunit_both(m, U, C, dslot);
}

// md::force_gradient_integrator (599-661): fine fields only; dt is the
// reference's Farg.dt (its dt argument is unused there).
int fg_md(const Multi& m, void* U, void* P, void* C, const void* xi, const mag_action& act, int steps, double dt,
          int dslot) {
  int r;
  if ((r = m.wf->need_md())) return r;
  void *Ff = m.wf->F, *FgA = m.wf->Uaux;
  const double alpha = (3. - std::sqrt(3.)) * dt / 6.;
  const double beta = dt / std::sqrt(3.);
  const double gamma = (2. - std::sqrt(3.)) * dt * dt / 12.;
  if ((r = fine_update(m, U, P, alpha))) return r;
  for (int i = 0; i < steps; ++i) {
    for (int h = 0; h < 2; ++h) {
      if ((r = fforce(m, Ff, U, C, xi, act))) return r;
      if ((r = fine_update(m, FgA, Ff, gamma, U))) return r;  // FgA = e^{i gamma Ff} U
      if ((r = fforce(m, Ff, FgA, C, xi, act))) return r;
      if ((r = mom_update(m.c, m.gf, MAG_F64, P, Ff, 0.5 * dt))) return r;
      if (h == 0 && (r = fine_update(m, U, P, beta))) return r;
    }
    if ((r = fine_update(m, U, P, i < steps - 1 ? 2. * alpha : alpha))) return r;
  }
  return reunit_both(m, U, C, dslot);
}

// nested_integrator (663-695)
int nested(const Multi& m, void* U, void* P, void* C, const void* xi, const mag_action& act, int steps, double dt,
           int dslot) {
  int r;
  if ((r = m.wc->need_md())) return r;
  void* Cm = m.wc->F;
  if ((r = cforce(m, Cm, U, P)) || (r = coarse_update(m, C, Cm, dt / 2.))) return r;
  for (int i = 0; i < steps; ++i) {
    if ((r = fg_md(m, U, P, C, xi, act, 1, dt, dslot))) return r;
    if ((r = cforce(m, Cm, U, P)) || (r = coarse_update(m, C, Cm, i < steps - 1 ? dt : dt / 2.))) return r;
  }
  return MAG_OK;
}

// md::get_hamiltonian_multi (74-114) into slots: K (s), plaq (s+1), rect (s+2), xi sum (s+3)
int ham_sums(const Multi& m, void* U, void* P, void* C, const void* xi, const mag_action& act, int s) {
  int r;
  if ((r = kinetic_sum(m.c, m.gf, MAG_F64, P, 0, s))) return r;
  if ((r = Ops<double, 18>::action_sums(m.c, m.gf, U, act.type == MAG_IWASAKI, s + 1))) return r;
  return xi_sum(m, U, C, xi, s + 3);
}
void ham_parts(const Multi& m, const mag_action& act, const double* h, double parts[3]) {
  parts[0] = h[0] / 2.;
  parts[1] = action_from_sums(act, h[1], act.type == MAG_IWASAKI ? h[2] : 0.);
  parts[2] = -h[3] * (1. + m.mp.md_alpha);
}

int check_act(mag_ctx c, const mag_action* act) {
  if (!act) return MAG_ERR_ARG;
  if (act->n_stout > 0) return fail(c, MAG_ERR_UNSUPPORTED, "the coarse/fine HMC uses the thin gauge action");
  return MAG_OK;
}

}  // namespace

extern "C" {

int mag_xi_energy(mag_field UF, mag_field UC, mag_field xi, const mag_multi_params* mp, double* out) {
  Multi m;
  int r;
  if ((r = make_multi(UF, mp, 0, &m)) || (r = check_coarse(m, UC, f64link)) || (r = check_coarse(m, xi, f64real)))
    return r;
  if ((r = xi_sum(m, UF->p, UC->p, xi->p, 0)) || (r = fetch_scalars(m.c, 1))) return r;
  if (out) *out = -m.c->h_scalar[0] * (1. + mp->md_alpha);
  return MAG_OK;
}

int mag_multi_gb(mag_field Gb, mag_field UF, const mag_multi_params* mp) {
  Multi m;
  int r;
  if ((r = make_multi(UF, mp, 0, &m)) || (r = check_coarse(m, Gb, f64link))) return r;
  MultiArgs a = args(m, UF->p, nullptr, nullptr);
  a.out = LV{(double*)Gb->p, m.gc.V};
  return mlaunch(m, "multi_gb", k_multi_gb, cgrid(m, 4), a);
}

int mag_multi_get_Q(mag_field Q, mag_field UF, const mag_multi_params* mp) {
  Multi m;
  int r;
  if ((r = make_multi(UF, mp, 0, &m)) || (r = check_coarse(m, Q, f64alg))) return r;
  MultiArgs a = args(m, UF->p, nullptr, nullptr);
  a.F = AV{(double*)Q->p, m.gc.V};
  return mlaunch(m, "multi_getq", k_multi_getq, cgrid(m, 4), a);
}

int mag_multi_fforce(mag_field F, mag_field UF, mag_field UC, mag_field xi, const mag_action* act,
                     const mag_multi_params* mp) {
  Multi m;
  int r;
  if ((r = make_multi(UF, mp, 0, &m)) || (r = check_act(m.c, act)) || (r = check_fine(m, F, f64alg)) ||
      (r = check_coarse(m, UC, f64link)) || (r = check_coarse(m, xi, f64real)))
    return r;
  return fforce(m, F->p, UF->p, UC->p, xi->p, *act);
}

int mag_multi_cforce(mag_field CF, mag_field UF, mag_field PF, const mag_multi_params* mp) {
  Multi m;
  int r;
  if ((r = make_multi(UF, mp, 0, &m)) || (r = check_fine(m, PF, f64alg)) || (r = check_coarse(m, CF, f64alg)))
    return r;
  return cforce(m, CF->p, UF->p, PF->p);
}

int mag_multi_heatbath(mag_field UC, mag_field UF, const mag_multi_params* mp, uint64_t seed, uint64_t traj) {
  Multi m;
  int r;
  if ((r = make_multi(UF, mp, 0, &m)) || (r = check_coarse(m, UC, f64link))) return r;
  MultiArgs a = args(m, UF->p, UC->p, nullptr);
  a.seed = seed;
  a.traj = traj;
  return mlaunch(m, "multi_heatbath", k_multi_heatbath, cgrid(m, 4), a);
}

int mag_multi_zeta_T(mag_field UF, mag_field PF, mag_field UC, double dt, int mag, const mag_multi_params* mp) {
  Multi m;
  int r;
  if ((r = make_multi(UF, mp, mag, &m)) || (r = check_fine(m, PF, f64alg)) || (r = check_coarse(m, UC, f64link)))
    return r;
  if ((r = m.wc->need_md())) return r;
  return zeta_T(m, UF->p, PF->p, UC->p, m.wc->F, dt);
}

int mag_multi_integrate(mag_field UF, mag_field PF, mag_field UC, mag_field xi, const mag_action* act,
                        const mag_multi_params* mp, int which, int steps, double dt, double max_dev[2]) {
  Multi m;
  int r;
  if (!act || steps < 1) return MAG_ERR_ARG;
  if ((r = make_multi(UF, mp, act->mag, &m)) || (r = check_act(m.c, act)) || (r = check_fine(m, PF, f64alg)) ||
      (r = check_coarse(m, UC, f64link)) || (r = check_coarse(m, xi, f64real)))
    return r;
  switch (which) {
    case MAG_MULTI_ZETA_H: r = zeta_H(m, UF->p, PF->p, UC->p, xi->p, *act, steps, dt, 0); break;
    case MAG_MULTI_FG: r = fg_md(m, UF->p, PF->p, UC->p, xi->p, *act, steps, dt, 0); break;
    case MAG_MULTI_NESTED: r = nested(m, UF->p, PF->p, UC->p, xi->p, *act, steps, dt, 0); break;
    default: return fail(m.c, MAG_ERR_ARG, "unknown coarse/fine integrator");
  }
  if (r || (r = fetch_scalars(m.c, 2))) return r;
  if (max_dev) {
    max_dev[0] = m.c->h_scalar[0];
    max_dev[1] = m.c->h_scalar[1];
  }
  return MAG_OK;
}

int mag_multi_hamiltonian(mag_field UF, mag_field PF, mag_field UC, mag_field xi, const mag_action* act,
                          const mag_multi_params* mp, double parts[3], double* H) {
  Multi m;
  int r;
  if ((r = make_multi(UF, mp, 0, &m)) || (r = check_act(m.c, act)) || (r = check_fine(m, PF, f64alg)) ||
      (r = check_coarse(m, UC, f64link)) || (r = check_coarse(m, xi, f64real)))
    return r;
  if ((r = ham_sums(m, UF->p, PF->p, UC->p, xi->p, *act, 0)) || (r = fetch_scalars(m.c, 4))) return r;
  double p[3];
  ham_parts(m, *act, m.c->h_scalar, p);
  if (parts) std::memcpy(parts, p, sizeof p);
  if (H) *H = p[0] + p[1] + p[2];
  return MAG_OK;
}

// ------------------------------------------------------------ run_hmc_multi
int mag_multi_create(mag_ctx ctx, const int fine_dims[4], const mag_action* act, const mag_hmc_params* p,
                     const mag_multi_params* mp, mag_multi* out) {
  if (!ctx || !fine_dims || !act || !p || !mp || !out) return MAG_ERR_ARG;
  *out = nullptr;
  int r;
  if ((r = check_act(ctx, act))) return r;
  if (p->precision != MAG_F64) return fail(ctx, MAG_ERR_UNSUPPORTED, "the coarse/fine HMC runs in FP64");
  if (p->trajectory_length < 1) return fail(ctx, MAG_ERR_ARG, "trajectory_length must be >= 1");
  int dc[4];
  for (int i = 0; i < 4; ++i) {
    if (fine_dims[i] % 2) return fail(ctx, MAG_ERR_GEOMETRY, "fine extents of the coarse/fine HMC must be even");
    dc[i] = fine_dims[i] / 2;
  }
  auto* m = new mag_multi_s();
  m->ctx = ctx;
  std::memcpy(m->fdims, fine_dims, sizeof m->fdims);
  m->act = *act;
  m->p = *p;
  m->mp = *mp;
  if ((r = mag_field_create(ctx, fine_dims, MAG_LINK18, MAG_F64, &m->UF)) ||
      (r = mag_field_create(ctx, fine_dims, MAG_LINK18, MAG_F64, &m->UFb)) ||
      (r = mag_field_create(ctx, fine_dims, MAG_ALG8, MAG_F64, &m->PF)) ||
      (r = mag_field_create(ctx, dc, MAG_LINK18, MAG_F64, &m->UC)) ||
      (r = mag_field_create(ctx, dc, MAG_LINK18, MAG_F64, &m->UCb)) ||
      (r = mag_field_create(ctx, dc, MAG_REAL1, MAG_F64, &m->xi)) || (r = mag_unit_links(m->UF)) ||
      (r = mag_unit_links(m->UC))) {
    mag_multi_destroy(m);
    return r;
  }
  *out = m;
  return MAG_OK;
}

int mag_multi_destroy(mag_multi m) {
  if (!m) return MAG_ERR_ARG;
  for (mag_field f : {m->UF, m->UFb, m->PF, m->UC, m->UCb, m->xi})
    if (f) mag_field_destroy(f);
  delete m;
  return MAG_OK;
}

int mag_multi_fields(mag_multi m, mag_field* UF, mag_field* UC) {
  if (!m) return MAG_ERR_ARG;
  if (UF) *UF = m->UF;
  if (UC) *UC = m->UC;
  return MAG_OK;
}

int mag_multi_run(mag_multi mc, int ntraj, mag_multi_result* results) {
  if (!mc || ntraj < 0) return MAG_ERR_ARG;
  Multi m;
  int r;
  if ((r = make_multi(mc->UF, &mc->mp, mc->act.mag, &m))) return r;
  mag_ctx c = mc->ctx;
  const mag_action& act = mc->act;
  // the state at entry is the accepted state (FgField_ext, CgField_ext)
  if ((r = copy_dev(c, mc->UFb->p, mc->UF->p, mc->UF->bytes)) ||
      (r = copy_dev(c, mc->UCb->p, mc->UC->p, mc->UC->bytes)))
    return r;
  for (int t = 0; t < ntraj; ++t) {
    const long i = mc->traj;
    void *U = mc->UF->p, *P = mc->PF->p, *C = mc->UC->p, *X = mc->xi->p;
    // init_momentum (878), heatbath (881-884), init_xi (893, xi = XI0)
    if ((r = mom_refresh(c, m.gf, MAG_F64, P, mc->p.seed, (uint64_t)i, 1.))) return r;
    MultiArgs ha = args(m, U, C, nullptr);
    ha.seed = mc->p.seed;
    ha.traj = (uint64_t)i;
    if ((r = mlaunch(m, "multi_heatbath", k_multi_heatbath, cgrid(m, 4), ha))) return r;
    if ((r = fill_dev(c, (double*)X, mc->mp.xi0, m.gc.V * 4))) return r;
    // H_old, zeta_H, H_new (896-901)
    if ((r = ham_sums(m, U, P, C, X, act, 0))) return r;
    if ((r = zeta_H(m, U, P, C, X, act, mc->p.trajectory_length, mc->p.dt, 8))) return r;
    if ((r = ham_sums(m, U, P, C, X, act, 4))) return r;
    if ((r = fetch_scalars(c, 10))) return r;
    double po[3], pn[3];
    ham_parts(m, act, c->h_scalar, po);
    ham_parts(m, act, c->h_scalar + 4, pn);
    const double dev[2] = {c->h_scalar[8], c->h_scalar[9]};
    const double h_old = po[0] + po[1] + po[2], h_new = pn[0] + pn[1] + pn[2];
    const double die = host_uniform(mc->p.seed, RNG_ACCEPT, (uint64_t)i);
    bool acc = die < std::exp(h_old - h_new);
    const bool bad = !std::isfinite(h_old) || !std::isfinite(h_new);
    if (i < mc->p.num_forced_accept && !bad) acc = true;
    if (bad) acc = false;
    // accept: the new state becomes the backup; reject: restore both fields
    // (which also drops this trajectory's heat-bath update, 903-928)
    if (acc) {
      if ((r = copy_dev(c, mc->UFb->p, U, mc->UF->bytes)) || (r = copy_dev(c, mc->UCb->p, C, mc->UC->bytes)))
        return r;
    } else {
      if ((r = copy_dev(c, U, mc->UFb->p, mc->UF->bytes)) || (r = copy_dev(c, C, mc->UCb->p, mc->UC->bytes)))
        return r;
    }
    ++mc->traj;
    if (results) {
      mag_multi_result& o = results[t];
      o.h_old = h_old;
      o.h_new = h_new;
      o.dh = h_new - h_old;
      o.die_roll = die;
      o.accept = acc ? 1 : 0;
      for (int k = 0; k < 3; ++k) {
        o.parts_old[k] = po[k];
        o.parts_new[k] = pn[k];
      }
      o.reunit_dev[0] = dev[0];
      o.reunit_dev[1] = dev[1];
      o.status = bad ? MAG_ERR_NAN : MAG_OK;
      if ((r = Ops<double, 18>::action_sums(c, m.gf, U, false, 0)) ||
          (r = Ops<double, 18>::action_sums(c, m.gc, C, false, 2)) || (r = fetch_scalars(c, 4)))
        return r;
      o.avg_plaq_fine = c->h_scalar[0] / (18. * m.gf.V);
      o.avg_plaq_coarse = c->h_scalar[2] / (18. * m.gc.V);
    }
  }
  return MAG_OK;
}

}  // extern "C"
