// gfx950 kernels of the constrained-HMC path. One thread per link (x, mu);
// a block covers 256 consecutive sites of one direction, so mu and every
// lattice extent/stride derived from it are wavefront-uniform (scalar
// registers); only the site coordinates are per-lane.
//
// Streaming kernels take blockIdx.x = site block, blockIdx.y = mu. Stencil
// kernels (staples, stout levels) take a 1D grid of 4 * nb tiles (one per
// 256-site block and direction), remapped by stencil_tile() so that each of
// the 8 XCDs (blocks go to XCDs round-robin) works on its own slab of the
// lattice, the four directions of a site block are dispatched back to back,
// and the neighbours of the tiles in flight are in flight or just done: the
// links they share are then fetched into that XCD's L2 once.
//
// Template parameters: TS storage real, NC reals per stored link (12 or 18),
// TC compute real. FP64 kernels (TS = TC = double, NC = 18) follow the
// reference arithmetic; the FP32 molecular-dynamics kernels use TS = TC =
// float with NC = 12. Hamiltonian reductions always accumulate in FP64.
#pragma once
#include <type_traits>

#include "cayley.hpp"
#include "lattice.hpp"
#include "philox.hpp"

namespace mag {

constexpr int BLOCK = 256;

// The FP32 stencil kernels need ~150-170 VGPRs. The waves per SIMD follow
// from the VGPR budget (512 per lane): <= 168 gives 3 waves, <= 256 gives 2.
// FP32 instances are held to 3 waves so that none falls to 2 (which cost
// the Iwasaki top kernel 20 %). FP64 stout kernels are held to 2 waves:
// unbounded they take 256 VGPRs plus 4-20 AGPRs and run 1 wave; at 2 waves
// they spill 20-116 bytes per lane and the backward and top kernels ran 16 %
// and 9 % faster. The FP64 thin-link force runs at 1 wave without spills: at
// 2 it spilled 116 bytes (free links) or 536 (constraint classes) and was 6 %
// slower on the grouped layout (the top kernel, with the same staple sum,
// runs equally fast at 1 and 2 waves there).
template <class TC>
constexpr int stencil_waves() { return sizeof(TC) == sizeof(float) ? 3 : 2; }
template <class TC, bool CONS>
constexpr int gforce_waves() { return sizeof(TC) == sizeof(float) ? 3 : 1; }
#define MAG_STENCIL_OCC __attribute__((amdgpu_waves_per_eu(stencil_waves<TC>())))
#define MAG_GFORCE_OCC __attribute__((amdgpu_waves_per_eu(gforce_waves<TC, CONS>())))

// Blocks are dispatched round-robin over the 8 XCDs (XCD = linear block id
// mod 8). When L3 splits into 8 slabs and an x3 slice is a whole number of
// blocks, XCD k owns x3 in [k L3/8, (k+1) L3/8) and walks it with the
// direction fastest, then x3, then the blocks of the slice in site order
// (for a 32^4 lattice: the (x0, x1) block, then x2): the tiles in flight on
// one XCD (about 100) then cover whole (x0, x1, x3) slabs at one or two
// values of x2, so their x3 neighbours are in flight and their x2 neighbours
// were just done. X2_INNER (needs the (x0, x1) plane to be a whole number of
// blocks) takes x2 before the (x0, x1) block: the FP64 top and backward
// kernels, whose 144-byte links make two x2 layers of whole slabs overflow
// L2, then find their x2 neighbours among the tiles of one block column
// (FP64 trajectory 2.7 % faster; the forward kernel and FP32 keep the first
// order). Otherwise XCD k takes tiles [k nb/8, (k+1) nb/8) in lexicographic
// order. nb = blocks in the launch, b = this block; every (mu, sblk) of the
// 4 V / BLOCK tiles comes out of exactly one b.
template <bool X2_INNER = false>
MAG_HD void tile_map(const Geom& g, int nb, int b, int& mu, int& sblk) {
  const int plane = g.L[0] * g.L[1];
  const int slice = plane * g.L[2];
  const bool slab = nb % 8 == 0 && g.L[3] % 8 == 0 && nb == 4 * (g.V / BLOCK) &&
                    (X2_INNER ? plane % BLOCK == 0 : slice % BLOCK == 0);
  if (slab) {
    const int xcd = b % 8, j = b / 8;
    const int h3 = g.L[3] / 8;
    mu = j & 3;
    int r = j >> 2;
    const int dx3 = r % h3;
    r /= h3;
    int bs;  // block within the x3 slice
    if (X2_INNER) {
      const int nb01 = plane / BLOCK;
      const int x2 = r % g.L[2];
      bs = x2 * nb01 + r / g.L[2];
    } else {
      bs = r;
    }
    const int x3 = xcd * h3 + dx3;
    sblk = x3 * (slice / BLOCK) + bs;
    return;
  }
  const int t = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
  mu = t & 3;
  sblk = t >> 2;
}
template <bool X2_INNER = false>
MAG_D void stencil_tile(const Geom& g, int& mu, int& sblk) {
  tile_map<X2_INNER>(g, gridDim.x, blockIdx.x, mu, sblk);
}

// ======================================================= staples and forces
enum ActionKind : int { ACT_WILSON = 0, ACT_IWASAKI = 1 };

// Site-index navigation of the staple sum. GNav works on global site
// indices; TNav (the LDS-tiled top kernel) pairs each with its index in the
// block's tile of (x0, x1) rows and a wavefront-uniform flag that says
// whether the displacement stays in that tile.
struct GNav {
  const Geom& g;
  const Site& p;
  MAG_D int at() const { return p.s; }
  MAG_D int off(int mu, int d) const { return off_rt(g, p, mu, d); }
};

// A view of an SU(3) link field whose loads read 18-real links as rows 0 and
// 1 and rebuild row 2 (two thirds of the groups).
template <class TS, int NC>
struct SU3Read {
  LinkView<TS, NC> u;
  template <class TC>
  MAG_D M3<TC> load(int mu, int s) const { return u.template load_su3<TC, true>(mu, s); }
  template <class TC>
  MAG_D M3<TC> load_r2(int mu, int s) const { return u.template load_r2<TC>(mu, s); }
};

// Undaggered staple sum S of link (x, mu): c0 * (6 plaquette staples)
// + c1 * (18 rectangle staples), the paths of get_staple_dagger
// (field-hmc.h:226-240) and get_rectangular_dagger (field-hmc.h:137-201),
// with c0 = 1 - 8 c1 (field-hmc.h:262-269). For Wilson, c0 = 1 and no
// rectangles. The reference's Sigma^+ is S^+.
//
// Rectangles share sub-products per plane nu: the three paths that end in
// U_nu(x+mu)^+ (plaquette, 1x2 tall, 2x1 backward) are summed before the
// last multiply, as are the three that end in U_nu(x-nu+mu); the 2x1 forward
// rectangles of every plane share the trailing U_mu(x+mu)^+, applied once
// after the nu loop. 62 instead of 85 3x3 products per link. Every partial
// path is a product of links, so it is formed in two rows (mul_*_r2) and
// completed once (reconstruct_row2) before it enters a sum.
//
// The terms of staple_sum in the plane (mu, nu): S gets the plaquette
// staples and the tall and backward rectangles, R the forward rectangles
// (multiplied by U_mu(x+mu)^+ once, after the last plane).
template <int ACT, class TC, class UV, class NAV, class I>
MAG_D void staple_plane(const UV& u, const NAV& nv, int mu, int nu, const I& xpm, const I& xmm, const I& xp2m, TC c0,
                        TC c1, M3<TC>& S, M3<TC>& R) {
  const auto opn = nv.off(nu, 1), omn = nv.off(nu, -1);
  const auto xpn = nv.at() + opn, xmn = nv.at() + omn;
  if (ACT == ACT_WILSON) {
    // U_nu(x) U_mu(x+nu) U_nu(x+mu)^+
    add_su3_r2(S, mul_nd_r2(mul_r2(u.template load_r2<TC>(nu, nv.at()), u.template load<TC>(mu, xpn)),
                            u.template load<TC>(nu, xpm)));
    // U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu)
    add_su3_r2(S, mul_r2(mul_dn_r2(u.template load<TC>(nu, xmn), u.template load<TC>(mu, xmn)),
                         u.template load<TC>(nu, xpm + omn)));
  } else {
    const auto op2n = nv.off(nu, 2), om2n = nv.off(nu, -2);
    // ---- upper side of the plane: the paths that end in U_nu(x+mu)^+
    //   G = c0 U_nu(x) U_mu(x+nu)                                (plaquette)
    //     + c1 U_nu(x) U_nu(x+nu) U_mu(x+2nu) U_nu(x+mu+nu)^+    (tall rectangle)
    //     + c1 U_mu(x-mu)^+ U_nu(x-mu) U_mu(x-mu+nu) U_mu(x+nu)  (backward rectangle)
    // each an SU(3) product formed in two rows; S += G U_nu(x+mu)^+.
    {
      const M3<TC> Un_x = u.template load_r2<TC>(nu, nv.at());
      const M3<TC> Um_xpn = u.template load<TC>(mu, xpn);
      const M3<TC> P = mul_r2(Un_x, Um_xpn);
      // forward rectangle R1' = P U_mu(x+nu+mu) U_nu(x+2mu)^+
      add_su3_r2(R, mul_nd_r2(mul_r2(P, u.template load<TC>(mu, xpm + opn)), u.template load<TC>(nu, xp2m)));
      M3<TC> G = zero3<TC>();
      axpy_su3_r2(G, c0, P);
      {
        M3<TC> T = mul_r2(Un_x, u.template load<TC>(nu, xpn));
        T = mul_r2(T, u.template load<TC>(mu, nv.at() + op2n));
        axpy_su3_r2(G, c1, mul_nd_r2(T, u.template load<TC>(nu, xpm + opn)));
      }
      {
        M3<TC> T = mul_dn_r2(u.template load<TC>(mu, xmm), u.template load<TC>(nu, xmm));
        T = mul_r2(T, u.template load<TC>(mu, xmm + opn));
        axpy_su3_r2(G, c1, mul_r2(T, Um_xpn));
      }
      mul_nd_acc(S, G, u.template load<TC>(nu, xpm));
    }
    // ---- lower side: the paths that end in U_nu(x-nu+mu)
    //   G = c0 U_nu(x-nu)^+ U_mu(x-nu)
    //     + c1 U_nu(x-nu)^+ U_nu(x-2nu)^+ U_mu(x-2nu) U_nu(x-2nu+mu)
    //     + c1 U_mu(x-mu)^+ U_nu(x-mu-nu)^+ U_mu(x-mu-nu) U_mu(x-nu)
    {
      const M3<TC> Un_xmn = u.template load<TC>(nu, xmn);
      const M3<TC> Um_xmn = u.template load<TC>(mu, xmn);
      const M3<TC> P = mul_dn_r2(Un_xmn, Um_xmn);
      // forward rectangle R2' = P U_mu(x-nu+mu) U_nu(x-nu+2mu)
      add_su3_r2(R, mul_r2(mul_r2(P, u.template load<TC>(mu, xpm + omn)), u.template load<TC>(nu, xp2m + omn)));
      M3<TC> G = zero3<TC>();
      axpy_su3_r2(G, c0, P);
      {
        M3<TC> T = mul_dd_r2(Un_xmn, u.template load<TC>(nu, nv.at() + om2n));
        T = mul_r2(T, u.template load<TC>(mu, nv.at() + om2n));
        axpy_su3_r2(G, c1, mul_r2(T, u.template load<TC>(nu, xpm + om2n)));
      }
      {
        M3<TC> T = mul_dd_r2(u.template load<TC>(mu, xmm), u.template load<TC>(nu, xmm + omn));
        T = mul_r2(T, u.template load<TC>(mu, xmm + omn));
        axpy_su3_r2(G, c1, mul_r2(T, Um_xmn));
      }
      mul_acc(S, G, u.template load<TC>(nu, xpm + omn));
    }
  }
}

template <int ACT, class TC, class UV, class NAV>
MAG_D M3<TC> staple_sum_nav(const UV& u, const NAV& nv, int mu, TC c0, TC c1) {
  M3<TC> S = zero3<TC>();
  M3<TC> R = zero3<TC>();
  const auto xpm = nv.at() + nv.off(mu, 1);
  const auto xmm = nv.at() + nv.off(mu, -1);
  const auto xp2m = nv.at() + nv.off(mu, 2);
#pragma unroll 1
  for (int i = 0; i < 3; ++i) {
    const int nu = i + (i >= mu ? 1 : 0);
    staple_plane<ACT, TC>(u, nv, mu, nu, xpm, xmm, xp2m, c0, c1, S, R);
  }
  if (ACT == ACT_IWASAKI) {
    R = scale(R, c1);
    mul_nd_acc(S, R, u.template load<TC>(mu, xpm));
  }
  return S;
}

template <int ACT, class TS, int NC, class TC>
MAG_D M3<TC> staple_sum(const LinkView<TS, NC>& u, const Geom& g, const Site& p, int mu, TC c0, TC c1) {
  return staple_sum_nav<ACT, TC>(u, GNav{g, p}, mu, c0, c1);
}

// Output of a force evaluation (chosen at run time; wave-uniform).
enum ForceOut : int {
  FOUT_F = 0,     // store F (8 algebra components)
  FOUT_PADD = 1,  // P += eps F           (evolve_momentum, field-hmc.h:427-440)
  FOUT_UAUX = 2,  // Uaux = e^{i eps F} U (FG auxiliary step, field-hmc.h:546-547)
};

template <class TS, int NC>
struct ForceArgs {
  Geom g;
  LinkView<TS, NC> U;     // links the force is evaluated on
  AlgView<TS> F;          // FOUT_F
  AlgView<TS> P;          // FOUT_PADD
  LinkView<TS, NC> Uaux;  // FOUT_UAUX
  double beta, c1;
  int mag;
  int out;
  double eps;
};

template <class TS, int NC, class TC>
MAG_D void force_epilogue(const ForceArgs<TS, NC>& a, int mu, int s, const M3<TC>& Ux, TC f[8]) {
  if (a.out == FOUT_F) {
    a.F.store(mu, s, f);
  } else if (a.out == FOUT_PADD) {
    TC p[8];
    a.P.load_nt(mu, s, p);
    const TC e = TC(a.eps);
#pragma unroll
    for (int k = 0; k < 8; ++k) p[k] = mfma(e, f[k], p[k]);
    a.P.store_nt(mu, s, p);
  } else {
    const TC e = TC(a.eps);
#pragma unroll
    for (int k = 0; k < 8; ++k) f[k] *= e;
    const M3<TC> Q = alg_to_mat(f);
    a.Uaux.store_nt(mu, s, exp_left(Q, Ux));
  }
}

// get_force (field-hmc.h:339-425): F = (i beta/3) TA(W) with W = U S^+ for
// free links, W = U(x+mu) S(x+mu)^+ - S(x)^+ U(x) for constraint classes 1 and
// 10, and F = 0 for class 100. Stored as F_a = -(beta/3) Im Tr(T_a W).
// CONS = false (mag = 0) compiles the free-link form alone (one staple sum,
// no class loop, U(x) loaded after it).
template <int ACT, bool CONS, class TS, int NC, class TC, bool R12 = false>
MAG_D void k_gauge_force_body(const ForceArgs<TS, NC>& a) {
  int mu, sblk;
  stencil_tile<sizeof(TC) == sizeof(double)>(a.g, mu, sblk);
  const int s = sblk * BLOCK + threadIdx.x;
  if (s >= a.g.V) return;
  const TC c1 = (ACT == ACT_IWASAKI) ? TC(a.c1) : TC(0);
  const TC c0 = TC(1) - TC(8) * c1;
  const Site p = site_of(a.g, s);
  M3<TC> Ux, W = zero3<TC>();
  if (!CONS) {
    // U(x) is loaded after the staple sum, so it is not live across it
    M3<TC> S;
    if constexpr (R12) S = staple_sum_nav<ACT, TC>(SU3Read<TS, NC>{a.U}, GNav{a.g, p}, mu, c0, c1);
    else S = staple_sum<ACT, TS, NC, TC>(a.U, a.g, p, mu, c0, c1);
    Ux = a.U.template load_su3<TC, R12>(mu, s);
    W = mul_nd(Ux, S);
  } else {
    const int cls = link_class_rt(p, mu, a.mag);
    Ux = a.U.template load_su3<TC, R12>(mu, s);
    if (cls != 100) {
      // class 0: one staple at x; classes 1, 10: staples at x and x+mu
      const int nsite = (cls == 0) ? 1 : 2;
      const int ys = s + off_rt(a.g, p, mu, 1);
#pragma unroll 1
      for (int k = 0; k < nsite; ++k) {
        const Site q = (k == 0) ? p : site_of(a.g, ys);
        M3<TC> S;
        if constexpr (R12) S = staple_sum_nav<ACT, TC>(SU3Read<TS, NC>{a.U}, GNav{a.g, q}, mu, c0, c1);
        else S = staple_sum<ACT, TS, NC, TC>(a.U, a.g, q, mu, c0, c1);
        if (cls == 0) W = mul_nd(Ux, S);
        else if (k == 0) W = sub(W, mul_dn(S, Ux));
        else mul_nd_acc(W, a.U.template load_su3<TC, R12>(mu, ys), S);
      }
    }
  }
  TC f[8];
  force_alg_from(W, f);
  const TC fac = TC(-a.beta / 3.);
#pragma unroll
  for (int k = 0; k < 8; ++k) f[k] *= fac;
  force_epilogue<TS, NC, TC>(a, mu, s, Ux, f);
}
template <int ACT, bool CONS, class TS, int NC, class TC, bool R12 = false>
__global__ __launch_bounds__(BLOCK) MAG_GFORCE_OCC void k_gauge_force(ForceArgs<TS, NC> a) { k_gauge_force_body<ACT, CONS, TS, NC, TC, R12>(a); }
template <int ACT, bool CONS, class TS, int NC, class TC, bool R12 = false>
__global__ __launch_bounds__(BLOCK) MAG_GFORCE_OCC void k_gauge_force_b(const ForceArgs<TS, NC>* t) { k_gauge_force_body<ACT, CONS, TS, NC, TC, R12>(t[blockIdx.z]); }

// ============================================================ link updates
// evolve_gauge_field (field-hmc.h:442-482) with the constraint classes:
// 0: U <- e^{i eps P(x)} U; 10: U <- U e^{-i eps P(x)};
// 1: U <- e^{i eps P(x-mu)} U e^{-i eps P(x)}; 100: U <- e^{i eps P(x-mu)} U.
// Uout may alias Uin (each thread reads and writes only its own link).
template <class TS, int NC>
struct UpdateArgs {
  Geom g;
  LinkView<TS, NC> Uin, Uout;
  AlgView<TS> P;
  double eps;
  int mag;
};

// R12: read an 18-real link as its rows 0 and 1 and rebuild row 2 (for
// callers whose links are unitary to rounding; two thirds of the bytes).
template <class TS, int NC, class TC, bool R12 = false>
MAG_D void k_link_update_body(const UpdateArgs<TS, NC>& a) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= a.g.V) return;
  const int mu = blockIdx.y;
  const Site p = site_of(a.g, s);
  const int cls = link_class_rt(p, mu, a.mag);
  const TC e = TC(a.eps);
  M3<TC> U;
  if constexpr (R12) {
    U = a.Uin.template load_r2_nt<TC>(mu, s);
    reconstruct_row2(U);
  } else {
    U = a.Uin.template load_nt<TC>(mu, s);
  }
  TC h[8];
  if (cls == 0 || cls == 1 || cls == 10) {
    a.P.load_nt(mu, s, h);
    const TC sg = (cls == 0) ? e : -e;
#pragma unroll
    for (int k = 0; k < 8; ++k) h[k] *= sg;
    const M3<TC> Qh = alg_to_mat(h);
    U = (cls == 0) ? exp_left(Qh, U) : exp_right(U, Qh);
  }
  if (cls == 1 || cls == 100) {
    a.P.load(mu, s + off_rt(a.g, p, mu, -1), h);
#pragma unroll
    for (int k = 0; k < 8; ++k) h[k] *= e;
    U = exp_left(alg_to_mat(h), U);
  }
  a.Uout.store_nt(mu, s, U);
}
template <class TS, int NC, class TC, bool R12 = false>
__global__ __launch_bounds__(BLOCK) void k_link_update(UpdateArgs<TS, NC> a) { k_link_update_body<TS, NC, TC, R12>(a); }
template <class TS, int NC, class TC, bool R12 = false>
__global__ __launch_bounds__(BLOCK) void k_link_update_b(const UpdateArgs<TS, NC>* t) {
  k_link_update_body<TS, NC, TC, R12>(t[blockIdx.z]);
}

// P += eps m F (field-hmc.h:427-440), for the unfused path.
template <class TS>
struct MomUpdateArgs {
  AlgView<TS> P, F;
  int V;
  double eps;
};
template <class TS>
MAG_D void k_mom_update_body(const MomUpdateArgs<TS>& a) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= a.V) return;
  const int mu = blockIdx.y;
  double p[8], f[8];
  a.P.load(mu, s, p);
  a.F.load(mu, s, f);
#pragma unroll
  for (int k = 0; k < 8; ++k) p[k] = mfma(a.eps, f[k], p[k]);
  a.P.store(mu, s, p);
}
template <class TS>
__global__ __launch_bounds__(BLOCK) void k_mom_update(MomUpdateArgs<TS> a) { k_mom_update_body<TS>(a); }
template <class TS>
__global__ __launch_bounds__(BLOCK) void k_mom_update_b(const MomUpdateArgs<TS>* t) { k_mom_update_body<TS>(t[blockIdx.z]); }

// ============================================================ stout chain
// Forward level k (hep-lat/0311018 eqs. 1-3; get_Q, field-multi.h:32-39,
// with the 6-plaquette staple): C = rho S, Q = TA(S U^+) (-i rho),
// U_{k+1} = e^{iQ} U_k. Stores Q (8 components), C (3x3) and U_{k+1}.
template <class TSI, int NCI, class TSO, int NCO>
struct StoutFwdArgs {
  Geom g;
  LinkView<TSI, NCI> U;
  AlgView<TSO> Q;
  MatView<TSO> C;
  LinkView<TSO, NCO> Unext;
  double rho;
  int store_qc;  // 0 when only the smeared field is wanted (Hamiltonian)
};

template <class TSI, int NCI, class TSO, int NCO, class TC, bool R12 = false>
MAG_D void k_stout_fwd_body(const StoutFwdArgs<TSI, NCI, TSO, NCO>& a) {
  int mu, sblk;
  stencil_tile(a.g, mu, sblk);
  const int s = sblk * BLOCK + threadIdx.x;
  if (s >= a.g.V) return;
  const Site p = site_of(a.g, s);
  M3<TC> S;
  if constexpr (R12) S = staple_sum_nav<ACT_WILSON, TC>(SU3Read<TSI, NCI>{a.U}, GNav{a.g, p}, mu, TC(1), TC(0));
  else S = staple_sum<ACT_WILSON, TSI, NCI, TC>(a.U, a.g, p, mu, TC(1), TC(0));
  const M3<TC> Ux = a.U.template load_su3<TC, R12>(mu, s);
  TC q[8];
  force_alg_from(mul_nd(S, Ux), q);  // -i TA(S U^+)
  const TC rho = TC(a.rho);
#pragma unroll
  for (int k = 0; k < 8; ++k) q[k] *= rho;
  const M3<TC> Q = alg_to_mat(q);
  a.Unext.store_nt(mu, s, exp_left(Q, Ux));
  if (a.store_qc) {
    a.Q.store_nt(mu, s, q);
    a.C.store_nt(mu, s, scale(S, rho));
  }
}
template <class TSI, int NCI, class TSO, int NCO, class TC, bool R12 = false>
__global__ __launch_bounds__(BLOCK) MAG_STENCIL_OCC void k_stout_fwd(StoutFwdArgs<TSI, NCI, TSO, NCO> a) { k_stout_fwd_body<TSI, NCI, TSO, NCO, TC, R12>(a); }
template <class TSI, int NCI, class TSO, int NCO, class TC, bool R12 = false>
__global__ __launch_bounds__(BLOCK) MAG_STENCIL_OCC void k_stout_fwd_b(const StoutFwdArgs<TSI, NCI, TSO, NCO>* t) { k_stout_fwd_body<TSI, NCI, TSO, NCO, TC, R12>(t[blockIdx.z]); }

// Lambda step of level k for one link, given Z' = dS/dU_{k+1} there
// (compute_Lambda, field-multi.h:180-269, eqs. 72-73):
//   Lambda_k = HT(Gamma)   (stored as its 8 generator components),
//   W_k = Z' e^{iQ_k} + i C_k^+ Lambda_k   (the local part of Z_k, eq. 75).
template <class TS, int NC, class TC, bool R12 = false>
MAG_D void stout_lambda_link(int mu, int s, const M3<TC>& Zp, const LinkView<TS, NC>& U, const AlgView<TS>& Qv,
                             const MatView<TS>& Cv, const AlgView<TS>& Lout, const MatView<TS>& Wout) {
  const M3<TC> Uk = U.template load_su3_nt<TC, R12>(mu, s);
  const M3<TC> Q = Qv.template load_mat_nt<TC>(mu, s);
  const M3<TC> X = mul(Uk, Zp);
  TC lam[8];
  M3<TC> eQ;
  ch_lambda(Q, X, lam, eQ);
  Lout.store_nt(mu, s, lam);
  const M3<TC> Lm = alg_to_mat(lam);
  M3<TC> W = mul(Zp, eQ);
  // + i C^+ Lambda
  const M3<TC> CL = mul_dh(Cv.template load_nt<TC>(mu, s), Lm);
#pragma unroll
  for (int k = 0; k < 9; ++k) W.e[k] = W.e[k] + cpx<TC>{-CL.e[k].i, CL.e[k].r};
  Wout.store_nt(mu, s, W);
}

// Top of the chain: Z_n = -(beta/3) S(U_n)^+ (derivative of the gauge action
// on the smeared field), then the Lambda step of level n-1.
// NCN: reals per stored link of U_n (FP64 12-real chains store the top level
// as 18 reals: the Iwasaki staple sum then rebuilds no rows and spills less).
template <class TS, int NC, int NCN = NC>
struct StoutTopArgs {
  Geom g;
  LinkView<TS, NCN> Un;     // smeared links U_n
  LinkView<TS, NC> Ulo;     // U_{n-1}
  AlgView<TS> Qlo;
  MatView<TS> Clo;
  AlgView<TS> Lout;
  MatView<TS> Wout;
  double beta, c1, rho;
};

// R12: the staple links of U_n (and U_{n-1}) read as rows 0 and 1 with row 2
// rebuilt; such a kernel runs at one wave per SIMD (its spills then go to
// AGPRs, not scratch).
template <int ACT, class TS, int NC, class TC, int NCN, bool R12 = false>
MAG_D void k_stout_top_body(const StoutTopArgs<TS, NC, NCN>& a) {
  int mu, sblk;
  stencil_tile<sizeof(TC) == sizeof(double)>(a.g, mu, sblk);
  const int s = sblk * BLOCK + threadIdx.x;
  if (s >= a.g.V) return;
  const TC c1 = (ACT == ACT_IWASAKI) ? TC(a.c1) : TC(0);
  const TC c0 = TC(1) - TC(8) * c1;
  const Site p = site_of(a.g, s);
  M3<TC> S;
  if constexpr (R12) S = staple_sum_nav<ACT, TC>(SU3Read<TS, NCN>{a.Un}, GNav{a.g, p}, mu, c0, c1);
  else S = staple_sum<ACT, TS, NCN, TC>(a.Un, a.g, p, mu, c0, c1);
  const M3<TC> Z = scale(dag(S), TC(-a.beta / 3.));
  stout_lambda_link<TS, NC, TC, R12>(mu, s, Z, a.Ulo, a.Qlo, a.Clo, a.Lout, a.Wout);
}
template <class TC, bool R12>
constexpr int top_waves() { return R12 ? 1 : stencil_waves<TC>(); }
template <int ACT, class TS, int NC, class TC, int NCN = NC, bool R12 = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(top_waves<TC, R12>()))) void k_stout_top(
    StoutTopArgs<TS, NC, NCN> a) {
  k_stout_top_body<ACT, TS, NC, TC, NCN, R12>(a);
}
template <int ACT, class TS, int NC, class TC, int NCN = NC, bool R12 = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(top_waves<TC, R12>()))) void k_stout_top_b(
    const StoutTopArgs<TS, NC, NCN>* t) {
  k_stout_top_body<ACT, TS, NC, TC, NCN, R12>(t[blockIdx.z]);
}

// ---- LDS-tiled variant of the FP32 Iwasaki top kernel (MAG_TOP_LDS).
// A block's 256 sites are whole x0 rows of one (x2, x3) slice. The block
// first stages the x0 and x1 links of its rows and of two halo rows on
// either side (256 + 4 L0 sites, at most TILE_NT) in LDS, either as the 12
// stored reals or rebuilt to 18 (padded to 20, five 16-byte groups). The
// staple terms of the (x0, x1) plane (mu and nu both 0 or 1) then read all
// their links from LDS; the other planes read global memory.
constexpr int TILE_NT = 384;

// Shapes the tile covers: L0 a power of two <= 32, L1 a power of two, and
// whole-block slices.
MAG_HD bool top_tile_ok(const Geom& g) {
  const int L0 = g.L[0], L1 = g.L[1];
  return (L0 == 4 || L0 == 8 || L0 == 16 || L0 == 32) && L1 > 0 && (L1 & (L1 - 1)) == 0 && (L0 * L1) % BLOCK == 0;
}

// A site of the block's slice within two rows of its own is at tile index
// (s - tb) mod (L0 L1), tb = the global index of the first staged site
// (tile rows follow x1 cyclically, so the power-of-two mask also covers the
// rows that wrap round the lattice).
template <int NCL>
struct TileLinks {
  using VT = PackVec<float, 4>::type;
  static constexpr int NG = NCL / 4;
  const VT* lds;  // [direction 0, 1][group][TILE_NT]
  int tb, pm;     // first staged site, L0 L1 - 1
  template <class TC, bool FULL>
  MAG_D M3<TC> get(int dir, int s) const {
    const int t = (s - tb) & pm;
    constexpr int NR = (FULL && NCL == 20) ? 5 : 3;
    float r[4 * NR];
#pragma unroll
    for (int gi = 0; gi < NR; ++gi) {
      const VT v = lds[(dir * NG + gi) * TILE_NT + t];
#pragma unroll
      for (int k = 0; k < 4; ++k) r[4 * gi + k] = v[k];
    }
    M3<TC> m;
#pragma unroll
    for (int k = 0; k < 6; ++k) m.e[k] = {TC(r[2 * k]), TC(r[2 * k + 1])};
    if constexpr (FULL && NCL == 20) {
#pragma unroll
      for (int k = 6; k < 9; ++k) m.e[k] = {TC(r[2 * k]), TC(r[2 * k + 1])};
    } else if constexpr (FULL) {
      reconstruct_row2(m);
    } else {
#pragma unroll
      for (int k = 6; k < 9; ++k) m.e[k] = {TC(0), TC(0)};
    }
    return m;
  }
};

// The staged links as a link view (directions 0 and 1, sites of the tile).
template <int NCL>
struct TileView {
  TileLinks<NCL> t;
  template <class TC>
  MAG_D M3<TC> load(int dir, int s) const { return t.template get<TC, true>(dir, s); }
  template <class TC>
  MAG_D M3<TC> load_r2(int dir, int s) const { return t.template get<TC, false>(dir, s); }
};

template <int NCL>
MAG_D void k_stout_top_tile_body(const StoutTopArgs<float, 12, 12>& a) {
  using VT = PackVec<float, 4>::type;
  constexpr int NG = NCL / 4;
  __shared__ VT tile[2 * NG * TILE_NT];
  int mu, sblk;
  stencil_tile<false>(a.g, mu, sblk);
  const int L0 = a.g.L[0], L1 = a.g.L[1], plane = L0 * L1;
  const int s0 = sblk * BLOCK;
  const int sl = s0 - s0 % plane;  // first site of the (x2, x3) slice
  const int x1lo = (s0 - sl) / L0;
  const int nt = BLOCK + 4 * L0;
  for (int ti = threadIdx.x; ti < nt; ti += BLOCK) {
    const int r = ti / L0, x0 = ti - r * L0;
    int x1 = (x1lo - 2 + r) % L1;
    if (x1 < 0) x1 += L1;
    const unsigned gs = (unsigned)(sl + x1 * L0 + x0);
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      if constexpr (NCL == 12) {
#pragma unroll
        for (int gi = 0; gi < 3; ++gi)
          tile[(d * NG + gi) * TILE_NT + ti] = at(Packed<float, 12>::group(a.Un.p, a.Un.V, d, gi), gs);
      } else {
        const M3<float> m = a.Un.template load<float>(d, (int)gs);
#pragma unroll
        for (int gi = 0; gi < 5; ++gi) {
          VT v;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int c = 4 * gi + k;
            v[k] = c < 18 ? ((c & 1) ? m.e[c / 2].i : m.e[c / 2].r) : 0.f;
          }
          tile[(d * NG + gi) * TILE_NT + ti] = v;
        }
      }
    }
  }
  __syncthreads();
  const int s = s0 + (int)threadIdx.x;
  const float c1 = float(a.c1);
  const float c0 = 1.f - 8.f * c1;
  const Site p = site_of(a.g, s);
  const GNav nv{a.g, p};
  const TileView<NCL> tv{TileLinks<NCL>{tile, sl + (x1lo - 2) * L0, plane - 1}};
  // the (x0, x1) plane from LDS, the other planes from global memory
  M3<float> S = zero3<float>();
  M3<float> R = zero3<float>();
  const int xpm = s + off_rt(a.g, p, mu, 1), xmm = s + off_rt(a.g, p, mu, -1), xp2m = s + off_rt(a.g, p, mu, 2);
#pragma unroll 1
  for (int i = 0; i < 3; ++i) {
    const int nu = i + (i >= mu ? 1 : 0);
    if (mu < 2 && nu < 2) staple_plane<ACT_IWASAKI, float>(tv, nv, mu, nu, xpm, xmm, xp2m, c0, c1, S, R);
    else staple_plane<ACT_IWASAKI, float>(a.Un, nv, mu, nu, xpm, xmm, xp2m, c0, c1, S, R);
  }
  R = scale(R, c1);
  mul_nd_acc(S, R, a.Un.template load<float>(mu, xpm));
  const M3<float> Z = scale(dag(S), float(-a.beta / 3.));
  stout_lambda_link<float, 12, float>(mu, s, Z, a.Ulo, a.Qlo, a.Clo, a.Lout, a.Wout);
}
template <int NCL>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(NCL == 12 ? 3 : 2))) void k_stout_top_tile(
    StoutTopArgs<float, 12, 12> a) {
  k_stout_top_tile_body<NCL>(a);
}
template <int NCL>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(NCL == 12 ? 3 : 2))) void k_stout_top_tile_b(
    const StoutTopArgs<float, 12, 12>* t) {
  k_stout_top_tile_body<NCL>(t[blockIdx.z]);
}

// Backward level k (eq. 75): Z_k = W_k + sum over nu of the six
// Lambda-inserted staple terms. With Y = -i rho U^+ Lambda every Y sits next
// to the link it was built from, so the terms are written in Lambda (8 reals
// per link instead of 18) around the shared products
// D = U_nu(x+mu) U_mu(x+nu)^+ and B = U_mu(x-nu) U_nu(x-nu+mu):
//   upper: i rho [(Lam_nu(x+mu) D - D Lam_mu(x+nu)) U_nu(x)^+ - D U_nu(x)^+ Lam_nu(x)]
//   lower: i rho U_nu(x+mu-nu)^+ [U_mu(x-nu)^+ (Lam_nu - Lam_mu)(x-nu)
//                                 - Lam_nu(x+mu-nu) U_mu(x-nu)^+] U_nu(x-nu)
// 10 products per plane. Then either the Lambda step of level k-1
// (LAST = false) or, at k = 0, the thin-link force F_a = Im Tr(T_a U Z_0)
// with a force epilogue.
template <class TS, int NC>
struct StoutBwdArgs {
  Geom g;
  LinkView<TS, NC> U;   // U_k
  AlgView<TS> L;        // Lambda_k
  MatView<TS> W;        // local part of Z_k
  // k > 0: Lambda step of level k-1
  LinkView<TS, NC> Ulo;
  AlgView<TS> Qlo;
  MatView<TS> Clo;
  AlgView<TS> Lout;
  MatView<TS> Wout;
  double rho;
  // k == 0: force output
  ForceArgs<TS, NC> f;
  MatView<TS> Zout;  // constrained runs: Z_0 is stored for k_force_from_Z
};

// c -= a b
template <class T>
MAG_HD void mul_sub(M3<T>& c, const M3<T>& a, const M3<T>& b) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = c(i, j);
#pragma unroll
      for (int k = 0; k < 3; ++k) s = cfma(cpx<T>{-a(i, k).r, -a(i, k).i}, b(k, j), s);
      c(i, j) = s;
    }
}
// c -= a b^+
template <class T>
MAG_HD void mul_nd_sub(M3<T>& c, const M3<T>& a, const M3<T>& b) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cpx<T> s = c(i, j);
#pragma unroll
      for (int k = 0; k < 3; ++k) s = cfma(cpx<T>{-a(i, k).r, -a(i, k).i}, cconj(b(j, k)), s);
      c(i, j) = s;
    }
}

template <class TS, int NC, class TC, bool R12>
MAG_D M3<TC> stout_six_terms(const StoutBwdArgs<TS, NC>& a, const Site& p, int mu) {
  const Geom& g = a.g;
  M3<TC> acc = zero3<TC>();
  const int xpm = p.s + off_rt(g, p, mu, 1);
#pragma unroll 1
  for (int i = 0; i < 3; ++i) {
    const int nu = i + (i >= mu ? 1 : 0);
    const int opn = off_rt(g, p, nu, 1), omn = off_rt(g, p, nu, -1);
    const int xpn = p.s + opn, xmn = p.s + omn, xpmmn = xpm + omn;
    {
      // D and E = D U_nu(x)^+ are products of links: formed in two rows, completed once
      M3<TC> D = mul_nd_r2(a.U.template load_r2<TC>(nu, xpm), a.U.template load_su3<TC, R12>(mu, xpn));
      reconstruct_row2(D);
      M3<TC> G = mul_hn(a.L.template load_mat<TC>(nu, xpm), D);
      mul_sub_nh(G, D, a.L.template load_mat<TC>(mu, xpn));
      const M3<TC> Un_x = a.U.template load_su3<TC, R12>(nu, p.s);
      mul_nd_acc(acc, G, Un_x);
      M3<TC> E = mul_nd_r2(D, Un_x);
      reconstruct_row2(E);
      mul_sub_nh(acc, E, a.L.template load_mat<TC>(nu, p.s));
    }
    {
      const M3<TC> Um_xmn = a.U.template load_su3<TC, R12>(mu, xmn);
      TC ln[8], lm[8];
      a.L.load(nu, xmn, ln);
      a.L.load(mu, xmn, lm);
#pragma unroll
      for (int k = 0; k < 8; ++k) ln[k] -= lm[k];
      M3<TC> K = mul_dh(Um_xmn, alg_to_mat(ln));
      mul_hd_sub(K, a.L.template load_mat<TC>(nu, xpmmn), Um_xmn);
      const M3<TC> K2 = mul_dn(a.U.template load_su3<TC, R12>(nu, xpmmn), K);
      mul_acc(acc, K2, a.U.template load_su3<TC, R12>(nu, xmn));
    }
  }
  // Z = W + i rho acc
  M3<TC> Z = a.W.template load_nt<TC>(mu, p.s);
  const TC rho = TC(a.rho);
#pragma unroll
  for (int k = 0; k < 9; ++k) Z.e[k] = {mfma(-rho, acc.e[k].i, Z.e[k].r), mfma(rho, acc.e[k].r, Z.e[k].i)};
  return Z;
}

template <bool LAST, class TS, int NC, class TC, bool R12 = false>
MAG_D void k_stout_bwd_body(const StoutBwdArgs<TS, NC>& a) {
  int mu, sblk;
  stencil_tile<sizeof(TC) == sizeof(double)>(a.g, mu, sblk);
  const int s = sblk * BLOCK + threadIdx.x;
  if (s >= a.g.V) return;
  const Site p = site_of(a.g, s);
  const M3<TC> Z = stout_six_terms<TS, NC, TC, R12>(a, p, mu);
  if (!LAST) {
    stout_lambda_link<TS, NC, TC, R12>(mu, s, Z, a.Ulo, a.Qlo, a.Clo, a.Lout, a.Wout);
  } else if (a.f.mag) {
    a.Zout.store(mu, s, Z);
  } else {
    const M3<TC> Ux = a.U.template load_su3<TC, R12>(mu, s);
    TC f[8];
    force_alg_from(mul(Ux, Z), f);
    force_epilogue<TS, NC, TC>(a.f, mu, s, Ux, f);
  }
}
template <bool LAST, class TS, int NC, class TC, bool R12 = false>
__global__ __launch_bounds__(BLOCK) MAG_STENCIL_OCC void k_stout_bwd(StoutBwdArgs<TS, NC> a) { k_stout_bwd_body<LAST, TS, NC, TC, R12>(a); }
template <bool LAST, class TS, int NC, class TC, bool R12 = false>
__global__ __launch_bounds__(BLOCK) MAG_STENCIL_OCC void k_stout_bwd_b(const StoutBwdArgs<TS, NC>* t) { k_stout_bwd_body<LAST, TS, NC, TC, R12>(t[blockIdx.z]); }

// Constrained runs with smearing: F from a stored Z_0 with the classes of
// field-hmc.h:353-378: W = U Z (0), U(x+mu) Z(x+mu) - Z U (1, 10), 0 (100);
// F_a = Im Tr(T_a W).
template <class TS, int NC>
struct ForceZArgs {
  ForceArgs<TS, NC> f;
  MatView<TS> Z;
};
template <class TS, int NC, class TC>
MAG_D void k_force_from_Z_body(const ForceZArgs<TS, NC>& z) {
  const ForceArgs<TS, NC>& a = z.f;
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= a.g.V) return;
  const int mu = blockIdx.y;
  const Site p = site_of(a.g, s);
  const int cls = link_class_rt(p, mu, a.mag);
  M3<TC> W = zero3<TC>();
  const M3<TC> Ux = a.U.template load<TC>(mu, s);
  if (cls == 0) {
    W = mul(Ux, z.Z.template load<TC>(mu, s));
  } else if (cls != 100) {
    const int ys = s + off_rt(a.g, p, mu, 1);
    W = sub(mul(a.U.template load<TC>(mu, ys), z.Z.template load<TC>(mu, ys)), mul(z.Z.template load<TC>(mu, s), Ux));
  }
  TC f[8];
  force_alg_from(W, f);
  a.F.store(mu, s, f);
}
template <class TS, int NC, class TC>
__global__ __launch_bounds__(BLOCK) void k_force_from_Z(ForceZArgs<TS, NC> a) { k_force_from_Z_body<TS, NC, TC>(a); }
template <class TS, int NC, class TC>
__global__ __launch_bounds__(BLOCK) void k_force_from_Z_b(const ForceZArgs<TS, NC>* t) {
  k_force_from_Z_body<TS, NC, TC>(t[blockIdx.z]);
}

// ============================================================ RNG kernels
// init_momentum (field-hmc.h:646-673): P = sum_a T_a g_a, g_a ~ N(0, sqrt m),
// drawn from the Philox momentum stream of link (s*4+mu), epoch = trajectory
// (read from device memory when epoch_ptr is set, so a captured trajectory
// replays with a new epoch).
template <class TS>
struct MomArgs {
  AlgView<TS> P;
  int V;
  uint64_t seed, epoch;
  const uint64_t* epoch_ptr;
  double sig;
};
template <class TS>
MAG_D void k_mom_refresh_body(const MomArgs<TS>& a) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= a.V) return;
  const int mu = blockIdx.y;
  const uint64_t epoch = a.epoch_ptr ? *a.epoch_ptr : a.epoch;
  const uint32_t link = (uint32_t)s * 4u + (uint32_t)mu;
  double g[8];
#pragma unroll
  for (uint32_t d = 0; d < 4; ++d) rng_gauss2(rng_draw(a.seed, RNG_MOMENTUM, epoch, link, d), g[2 * d], g[2 * d + 1]);
#pragma unroll
  for (int k = 0; k < 8; ++k) g[k] *= a.sig;
  a.P.store(mu, s, g);
}
template <class TS>
__global__ __launch_bounds__(BLOCK) void k_mom_refresh(MomArgs<TS> a) { k_mom_refresh_body<TS>(a); }
template <class TS>
__global__ __launch_bounds__(BLOCK) void k_mom_refresh_b(const MomArgs<TS>* t) { k_mom_refresh_body<TS>(t[blockIdx.z]); }

// Haar-random SU(3) links (Mezzadri 2007): Gram-Schmidt of a complex Gaussian
// matrix, det phase divided out.
template <class TS, int NC>
__global__ __launch_bounds__(BLOCK) void k_haar(LinkView<TS, NC> U, int V, uint64_t seed) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= V) return;
  const int mu = blockIdx.y;
  const uint32_t link = (uint32_t)s * 4u + (uint32_t)mu;
  M3<double> z;
#pragma unroll
  for (uint32_t d = 0; d < 9; ++d) {
    double g0, g1;
    rng_gauss2(rng_draw(seed, RNG_HAAR, 0, link, d), g0, g1);
    z.e[d] = {g0, g1};
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int q = 0; q < r; ++q) {
      cpx<double> c = {0., 0.};
#pragma unroll
      for (int j = 0; j < 3; ++j) c = c + cconj(z(q, j)) * z(r, j);
#pragma unroll
      for (int j = 0; j < 3; ++j) z(r, j) = z(r, j) - c * z(q, j);
    }
    double n = 0.;
#pragma unroll
    for (int j = 0; j < 3; ++j) n += z(r, j).r * z(r, j).r + z(r, j).i * z(r, j).i;
    n = 1. / sqrt(n);
#pragma unroll
    for (int j = 0; j < 3; ++j) z(r, j) = n * z(r, j);
  }
  const cpx<double> d = z(0, 0) * (z(1, 1) * z(2, 2) - z(1, 2) * z(2, 1)) -
                        z(0, 1) * (z(1, 0) * z(2, 2) - z(1, 2) * z(2, 0)) +
                        z(0, 2) * (z(1, 0) * z(2, 1) - z(1, 1) * z(2, 0));
  const double phi = atan2(d.i, d.r);
  z = cscale(z, cpx<double>{cos(-phi / 3.), sin(-phi / 3.)});
  U.store(mu, s, z);
}

template <class TS, int NC>
__global__ __launch_bounds__(BLOCK) void k_unit(LinkView<TS, NC> U, int V) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= V) return;
  U.store(blockIdx.y, s, unit3<double>());
}

// ============================================================ reductions
// Block sum in a fixed order: wavefront shuffles, then wave partials in LDS.
// The result depends only on the values and the block shape, never on timing.
MAG_D double block_sum(double v) {
  __shared__ double red[BLOCK / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) t += red[w];
  }
  __syncthreads();
  return t;
}
MAG_D double block_max(double v) {
  __shared__ double red[BLOCK / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) t = fmax(t, red[w]);
  }
  __syncthreads();
  return t;
}

// Per-site plaquette and rectangle Re Tr sums in FP64 (get_plaq,
// field-matrix.h:202-221; get_rectangular, field-matrix.h:249-297), one
// partial pair per block: part[2b] = plaquettes, part[2b+1] = rectangles.
template <class TS, int NC>
struct ActionArgs {
  LinkView<TS, NC> U;
  Geom g;
  double* part;
  int rect;
};
template <class TS, int NC>
MAG_D void k_action_partial_body(const ActionArgs<TS, NC>& aa) {
  const LinkView<TS, NC>& U = aa.U;
  const Geom& g = aa.g;
  double* part = aa.part;
  const int rect = aa.rect;
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  double sp = 0., sr = 0.;
  if (s < g.V) {
    const Site p = site_of(g, s);
#pragma unroll 1
    for (int pl = 0; pl < 6; ++pl) {
      // planes (mu, nu) with nu < mu: (1,0) (2,0) (2,1) (3,0) (3,1) (3,2)
      const int mu = pl < 1 ? 1 : (pl < 3 ? 2 : 3);
      const int nu = pl - (mu * (mu - 1)) / 2;
      const int xpm = s + off_rt(g, p, mu, 1);
      const int xpn = s + off_rt(g, p, nu, 1);
      // U_mu(x) U_nu(x+mu) [U_nu(x) U_mu(x+nu)]^+
      const M3<double> Um = U.template load<double>(mu, s);
      const M3<double> Un = U.template load<double>(nu, s);
      const M3<double> m1 = mul(Um, U.template load<double>(nu, xpm));
      const M3<double> n1 = mul(Un, U.template load<double>(mu, xpn));
      sp += retrace_mul_nd(m1, n1);
      if (rect) {
        const int xpmpn = xpm + off_rt(g, p, nu, 1);
        // 1x2 (tall in nu): U_nu(x) U_nu(x+nu) U_mu(x+2nu) [U_mu(x) U_nu(x+mu) U_nu(x+mu+nu)]^+
        {
          const M3<double> a = mul(mul(Un, U.template load<double>(nu, xpn)),
                                   U.template load<double>(mu, s + off_rt(g, p, nu, 2)));
          const M3<double> b = mul(m1, U.template load<double>(nu, xpmpn));
          sr += retrace_mul_nd(a, b);
        }
        // 2x1 (long in mu): U_nu(x) U_mu(x+nu) U_mu(x+mu+nu) [U_mu(x) U_mu(x+mu) U_nu(x+2mu)]^+
        {
          const M3<double> a = mul(n1, U.template load<double>(mu, xpmpn));
          const M3<double> b = mul(mul(Um, U.template load<double>(mu, xpm)),
                                   U.template load<double>(nu, s + off_rt(g, p, mu, 2)));
          sr += retrace_mul_nd(a, b);
        }
      }
    }
  }
  const double tp = block_sum(sp);
  const double tr = block_sum(sr);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = tp;
    part[2 * blockIdx.x + 1] = tr;
  }
}
template <class TS, int NC>
__global__ __launch_bounds__(BLOCK) void k_action_partial(ActionArgs<TS, NC> a) { k_action_partial_body<TS, NC>(a); }
template <class TS, int NC>
__global__ __launch_bounds__(BLOCK) void k_action_partial_b(const ActionArgs<TS, NC>* t) {
  k_action_partial_body<TS, NC>(t[blockIdx.z]);
}

// Sum p_a^2 over links of class != 100 (field-hmc.h:596-629), FP64.
template <class TS>
struct KineticArgs {
  AlgView<TS> P;
  Geom g;
  int mag;
  double* part;
};
template <class TS>
MAG_D void k_kinetic_partial_body(const KineticArgs<TS>& a) {
  const AlgView<TS>& P = a.P;
  const Geom& g = a.g;
  const int mag = a.mag;
  double* part = a.part;
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  double acc = 0.;
  if (s < g.V) {
    const Site p = site_of(g, s);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      if (link_class_rt(p, mu, mag) == 100) continue;
      double h[8];
      P.load(mu, s, h);
#pragma unroll
      for (int k = 0; k < 8; ++k) acc = mfma(h[k], h[k], acc);
    }
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}
template <class TS>
__global__ __launch_bounds__(BLOCK) void k_kinetic_partial(KineticArgs<TS> a) { k_kinetic_partial_body<TS>(a); }
template <class TS>
__global__ __launch_bounds__(BLOCK) void k_kinetic_partial_b(const KineticArgs<TS>* t) {
  k_kinetic_partial_body<TS>(t[blockIdx.z]);
}

// reunitarize (field-matrix.h:83-100): project every link whose first two
// rows are not all zero, FP64 arithmetic; max Frobenius change per block.
template <class TS, int NC>
struct ReunitArgs {
  LinkView<TS, NC> U;
  int V;
  double* part;
};
template <class TS, int NC>
MAG_D void k_reunitarize_body(const ReunitArgs<TS, NC>& a) {
  const LinkView<TS, NC>& U = a.U;
  const int V = a.V;
  double* part = a.part;
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  const int mu = blockIdx.y;
  double dev = 0.;
  if (s < V) {
    M3<double> m = U.template load<double>(mu, s);
    const M3<double> old = m;
    bool all_zero = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) all_zero = all_zero && m.e[k].r == 0. && m.e[k].i == 0.;
    if (!all_zero) unitarize(m);
    double d = 0.;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double dr = m.e[k].r - old.e[k].r, di = m.e[k].i - old.e[k].i;
      d += dr * dr + di * di;
    }
    // the row-2 entries of a 12-real link are rebuilt on every load, so
    // they carry no stored deviation
    dev = sqrt(d);
    U.store(mu, s, m);
  }
  const double t = block_max(dev);
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = t;
}
template <class TS, int NC>
__global__ __launch_bounds__(BLOCK) void k_reunitarize(ReunitArgs<TS, NC> a) { k_reunitarize_body<TS, NC>(a); }
template <class TS, int NC>
__global__ __launch_bounds__(BLOCK) void k_reunitarize_b(const ReunitArgs<TS, NC>* t) {
  k_reunitarize_body<TS, NC>(t[blockIdx.z]);
}

// Wilson loops (field-wloop.h:26-98): Re Tr of the R x T loop for every
// permutation of r, per site, FP64.
struct WloopArgs {
  Geom g;
  int nperm;
  int r[6][3];
  int T;
};
template <class TS, int NC>
__global__ __launch_bounds__(BLOCK) void k_wilson_loop(LinkView<TS, NC> U, WloopArgs a, double* part) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  double acc = 0.;
  if (s < a.g.V) {
    for (int q = 0; q < a.nperm; ++q) {
      Site y = site_of(a.g, s);
      M3<double> m = unit3<double>();
      auto fwd = [&](int d) {
        m = mul(m, U.template load<double>(d, y.s));
        y.s += off_rt(a.g, y, d, 1);
        y.x[d] = (y.x[d] + 1 == a.g.L[d]) ? 0 : y.x[d] + 1;
      };
      auto bwd = [&](int d) {
        y.s += off_rt(a.g, y, d, -1);
        y.x[d] = (y.x[d] == 0) ? a.g.L[d] - 1 : y.x[d] - 1;
        m = mul_nd(m, U.template load<double>(d, y.s));
      };
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < a.r[q][i]; ++j) fwd(i);
      for (int j = 0; j < a.T; ++j) fwd(3);
      for (int i = 2; i >= 0; --i)
        for (int j = 0; j < a.r[q][i]; ++j) bwd(i);
      for (int j = 0; j < a.T; ++j) bwd(3);
      acc += retrace(m);
    }
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Final reduction of block partials, one block, fixed order: each thread
// folds a strided slice, then the block tree. OP 0 = sum, 1 = max.
struct ReduceArgs {
  const double* part;
  int n, stride;
  double* out;
};
template <int OP>
MAG_D void k_final_reduce_body(const ReduceArgs& a) {
  double acc = 0.;
  for (int i = threadIdx.x; i < a.n; i += BLOCK) {
    const double v = a.part[(size_t)i * a.stride];
    acc = OP ? fmax(acc, v) : acc + v;
  }
  const double t = OP ? block_max(acc) : block_sum(acc);
  if (threadIdx.x == 0) *a.out = t;
}
template <int OP>
__global__ __launch_bounds__(BLOCK) void k_final_reduce(ReduceArgs a) { k_final_reduce_body<OP>(a); }
template <int OP>
__global__ __launch_bounds__(BLOCK) void k_final_reduce_b(const ReduceArgs* t) { k_final_reduce_body<OP>(t[blockIdx.z]); }

// out[i] = sum_j w[i][j] s[j] for i < nout (<= 3): linear combinations of
// reduction results (e.g. kinetic, potential and H from the plaquette,
// rectangle and momentum sums), written to caller memory. One block of 64.
struct LinCombArgs {
  const double* s;
  double* out;
  int nout;
  double w[3][4];
};
template <int DUMMY>
__global__ void k_lincomb(LinCombArgs a) {
  const int i = threadIdx.x;
  if (i < a.nout) a.out[i] = a.w[i][0] * a.s[0] + a.w[i][1] * a.s[1] + a.w[i][2] * a.s[2] + a.w[i][3] * a.s[3];
}

// Device copy and fill (used instead of runtime memcpy / memset nodes so a
// trajectory can be batched over chains and captured in a graph).
struct CopyArgs {
  double2* dst;
  const double2* src;
  size_t n;  // 16-byte words
};
// One 16-byte word per thread (the launch covers n), non-temporal: 1.4x the
// rate of a grid-stride loop over 4,096 blocks.
MAG_D void k_copy_body(const CopyArgs& a) {
  const size_t stride = (size_t)gridDim.x * BLOCK;
  using V2 = PackVec<double, 2>::type;
  const V2* src = (const V2*)a.src;
  V2* dst = (V2*)a.dst;
  for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < a.n; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}
template <int DUMMY>
__global__ __launch_bounds__(BLOCK) void k_copy(CopyArgs a) { k_copy_body(a); }
template <int DUMMY>
__global__ __launch_bounds__(BLOCK) void k_copy_b(const CopyArgs* t) { k_copy_body(t[blockIdx.z]); }

struct FillArgs {
  double* p;
  double v;
  int n;
};
template <int DUMMY>
__global__ __launch_bounds__(BLOCK) void k_fill(FillArgs a) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < a.n) a.p[i] = a.v;
}
template <int DUMMY>
__global__ __launch_bounds__(BLOCK) void k_fill_b(const FillArgs* t) {
  const FillArgs& a = t[blockIdx.z];
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < a.n) a.p[i] = a.v;
}

// ============================================================ conversions
template <class TSI, int NCI, class TSO, int NCO>
__global__ __launch_bounds__(BLOCK) void k_convert_links(LinkView<TSI, NCI> in, LinkView<TSO, NCO> out, int V) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= V) return;
  const int mu = blockIdx.y;
  out.store(mu, s, in.template load<double>(mu, s));
}

// max over links of |A - B| (Frobenius), per-block partials
template <class TS, int NC>
__global__ __launch_bounds__(BLOCK) void k_max_diff(LinkView<TS, NC> A, LinkView<TS, NC> B, int V, double* part) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  const int mu = blockIdx.y;
  double dev = 0.;
  if (s < V) {
    const M3<double> a = A.template load<double>(mu, s), b = B.template load<double>(mu, s);
    double d = 0.;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double dr = a.e[k].r - b.e[k].r, di = a.e[k].i - b.e[k].i;
      d += dr * dr + di * di;
    }
    dev = sqrt(d);
  }
  const double t = block_max(dev);
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = t;
}

// P <- s P (momentum flip for reversibility checks)
template <class TS>
__global__ __launch_bounds__(BLOCK) void k_alg_scale(AlgView<TS> P, int V, double sc) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= V) return;
  const int mu = blockIdx.y;
  double h[8];
  P.load(mu, s, h);
#pragma unroll
  for (int k = 0; k < 8; ++k) h[k] *= sc;
  P.store(mu, s, h);
}

// Batched per-matrix test hooks (FP64 or FP32 arithmetic).
template <class TC>
__global__ __launch_bounds__(BLOCK) void k_expiQ_batch(const double* Q, double* out, long n) {
  const long i = (long)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  M3<TC> q;
#pragma unroll
  for (int k = 0; k < 9; ++k) q.e[k] = {TC(Q[18 * i + 2 * k]), TC(Q[18 * i + 2 * k + 1])};
  const M3<TC> e = expiQ(q);
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    out[18 * i + 2 * k] = double(e.e[k].r);
    out[18 * i + 2 * k + 1] = double(e.e[k].i);
  }
}

template <class TC>
__global__ __launch_bounds__(BLOCK) void k_dexpiq_batch(const double* Q, const double* DQ, double* out, long n) {
  const long i = (long)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  M3<TC> q, dq;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    q.e[k] = {TC(Q[18 * i + 2 * k]), TC(Q[18 * i + 2 * k + 1])};
    dq.e[k] = {TC(DQ[18 * i + 2 * k]), TC(DQ[18 * i + 2 * k + 1])};
  }
  M3<TC> e, de;
  dexpiq(q, dq, e, de);
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    out[18 * i + 2 * k] = double(de.e[k].r);
    out[18 * i + 2 * k + 1] = double(de.e[k].i);
  }
}

template <class TC>
__global__ __launch_bounds__(BLOCK) void k_lambda_batch(const double* Q, const double* Sg, const double* U,
                                                        double* out, long n) {
  const long i = (long)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  M3<TC> q, sg, u;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    q.e[k] = {TC(Q[18 * i + 2 * k]), TC(Q[18 * i + 2 * k + 1])};
    sg.e[k] = {TC(Sg[18 * i + 2 * k]), TC(Sg[18 * i + 2 * k + 1])};
    u.e[k] = {TC(U[18 * i + 2 * k]), TC(U[18 * i + 2 * k + 1])};
  }
  TC lam[8];
  M3<TC> eQ;
  ch_lambda(q, mul(u, sg), lam, eQ);
  const M3<TC> L = alg_to_mat(lam);
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    out[18 * i + 2 * k] = double(L.e[k].r);
    out[18 * i + 2 * k + 1] = double(L.e[k].i);
  }
}

}  // namespace mag
