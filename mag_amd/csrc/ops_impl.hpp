// Definitions of Ops<TS, NC>: the launch schedules of the force (thin or
// stout-smeared), link updates, integrators and reductions. Included only by
// the per-format instantiation units (inst_f64.hip, inst_f32.hip).
#pragma once
#include <algorithm>
#include <cmath>

#include "kernels.hpp"
#include "launch.hpp"
#include "runtime.hpp"

// Reals per staged link of the LDS-tiled FP32 Iwasaki top kernel (12 or 20;
// 0: the untiled kernel).
#ifndef MAG_TOP_LDS
#define MAG_TOP_LDS 0
#endif

// FP64 18-real links read as rows 0 and 1 with row 2 rebuilt, a bit per
// kernel: 1 stout forward, 2 stout backward, 4 last backward, 8 link update
// (on by default), 16 the top kernel at one wave per SIMD (off: as fast as
// its full reads at two waves), 32 the thin-link force (off: within 1 %).
#ifndef MAG_FP64_READ12
#define MAG_FP64_READ12 15
#endif

namespace mag {

// whether Ops<TS, NC> reads its links as two rows in those kernels
template <class TS, int NC>
constexpr bool ops_read12(int bit) { return (MAG_FP64_READ12 & bit) != 0 && std::is_same<TS, double>::value && NC == 18; }

template <class TS, int NC>
LinkView<TS, NC> lview(void* p, const Geom& g) { return LinkView<TS, NC>{(TS*)p, g.V}; }
template <class TS>
AlgView<TS> aview(void* p, const Geom& g) { return AlgView<TS>{(TS*)p, g.V}; }
template <class TS>
MatView<TS> mview(void* p, const Geom& g) { return MatView<TS>{(TS*)p, g.V}; }

template <class TS, int NC>
int Ops<TS, NC>::link_update(mag_ctx c, const Geom& g, void* Uin, void* Uout, void* P, double eps, int mag) {
  UpdateArgs<TS, NC> a{g, lview<TS, NC>(Uin, g), lview<TS, NC>(Uout, g), aview<TS>(P, g), eps, mag};
  constexpr bool R12 = ops_read12<TS, NC>(8);
  return launch(c, "link_update", k_link_update<TS, NC, TS, R12>, k_link_update_b<TS, NC, TS, R12>, grid_links(g), a);
}

// Force of the (possibly smeared) action on U with one of the epilogues
// OUT_F -> F, OUT_PADD -> P += eps F, OUT_UAUX -> Uaux = e^{i eps F} U.
// Thin action: one kernel (k_gauge_force). Smeared action (n levels):
// n forward kernels, the top kernel (action derivative + Lambda of level
// n-1), n-1 backward kernels each fused with the next Lambda step, and the
// last backward kernel fused with the force epilogue.
template <class TS, int NC>
int Ops<TS, NC>::force(mag_ctx c, const Geom& g, void* U, const mag_action& act, int out, void* F, void* P,
                       void* Uaux, double eps, Workspace* ws) {
  int r;
  // constrained runs cannot fuse the auxiliary link update (it needs F(x-mu))
  if (act.mag && out == OUT_UAUX) {
    if ((r = ws->need_md())) return r;
    if ((r = force(c, g, U, act, OUT_F, ws->F, nullptr, nullptr, 0., ws))) return r;
    return link_update(c, g, U, Uaux, ws->F, eps, act.mag);
  }
  ForceArgs<TS, NC> fa{};
  fa.g = g;
  fa.U = lview<TS, NC>(U, g);
  fa.F = aview<TS>(F, g);
  fa.P = aview<TS>(P, g);
  fa.Uaux = lview<TS, NC>(Uaux, g);
  fa.beta = act.beta;
  fa.c1 = act.c1;
  fa.mag = act.mag;
  fa.out = out;
  fa.eps = eps;
  const bool iw = act.type == MAG_IWASAKI;
  if (act.n_stout <= 0) {
    constexpr bool R12 = ops_read12<TS, NC>(32);
    if (iw && act.mag)
      return launch(c, "force_iwasaki", k_gauge_force<ACT_IWASAKI, true, TS, NC, TS, R12>,
                    k_gauge_force_b<ACT_IWASAKI, true, TS, NC, TS, R12>, grid_stencil(g), fa);
    if (iw)
      return launch(c, "force_iwasaki", k_gauge_force<ACT_IWASAKI, false, TS, NC, TS, R12>,
                    k_gauge_force_b<ACT_IWASAKI, false, TS, NC, TS, R12>, grid_stencil(g), fa);
    if (act.mag)
      return launch(c, "force_wilson", k_gauge_force<ACT_WILSON, true, TS, NC, TS, R12>,
                    k_gauge_force_b<ACT_WILSON, true, TS, NC, TS, R12>, grid_stencil(g), fa);
    return launch(c, "force_wilson", k_gauge_force<ACT_WILSON, false, TS, NC, TS, R12>,
                  k_gauge_force_b<ACT_WILSON, false, TS, NC, TS, R12>, grid_stencil(g), fa);
  }
  const int n = act.n_stout;
  constexpr bool R12 = ops_read12<TS, NC>(1), R12B = ops_read12<TS, NC>(2), R12L = ops_read12<TS, NC>(4);
  if ((r = ws->need_stout(n))) return r;
  // FP64 12-real chains keep the top level U_n as 18 reals (written by the
  // last forward level, read only by the top kernel's staple sum).
  constexpr int NCN = (std::is_same<TS, double>::value && NC == 12) ? 18 : NC;
  if (NCN != NC && (r = ws->need_top18())) return r;
  auto Ulev = [&](int k) { return k == 0 ? U : ws->Us[k - 1]; };
  for (int k = 0; k < n; ++k) {
    if (NCN != NC && k == n - 1) {
      StoutFwdArgs<TS, NC, TS, NCN> fw{g, lview<TS, NC>(Ulev(k), g), aview<TS>(ws->Qs[k], g),
                                       mview<TS>(ws->Cs[k], g), lview<TS, NCN>(ws->Utop, g), act.rho, 1};
      r = launch(c, "stout_fwd", k_stout_fwd<TS, NC, TS, NCN, TS>, k_stout_fwd_b<TS, NC, TS, NCN, TS>,
                 grid_stencil(g), fw);
    } else {
      StoutFwdArgs<TS, NC, TS, NC> fw{g, lview<TS, NC>(Ulev(k), g), aview<TS>(ws->Qs[k], g),
                                      mview<TS>(ws->Cs[k], g), lview<TS, NC>(ws->Us[k], g), act.rho, 1};
      r = launch(c, "stout_fwd", k_stout_fwd<TS, NC, TS, NC, TS, R12>, k_stout_fwd_b<TS, NC, TS, NC, TS, R12>,
                 grid_stencil(g), fw);
    }
    if (r) return r;
  }
  int cur = 0;
  {
    StoutTopArgs<TS, NC, NCN> tp{g,
                                 lview<TS, NCN>(NCN != NC ? ws->Utop : Ulev(n), g),
                                 lview<TS, NC>(Ulev(n - 1), g),
                                 aview<TS>(ws->Qs[n - 1], g),
                                 mview<TS>(ws->Cs[n - 1], g),
                                 aview<TS>(ws->L[cur], g),
                                 mview<TS>(ws->W, g),
                                 act.beta,
                                 act.c1,
                                 act.rho};
    bool tiled = false;  // the LDS-tiled FP32 Iwasaki top kernel, when compiled in and the shape fits
    if constexpr (MAG_TOP_LDS != 0 && std::is_same<TS, float>::value && NC == 12) {
      tiled = iw && top_tile_ok(g);
      if (tiled)
        r = launch(c, "stout_top", k_stout_top_tile<MAG_TOP_LDS>, k_stout_top_tile_b<MAG_TOP_LDS>, grid_stencil(g),
                   tp);
    }
    // FP64 (both link formats: the top level U_n is stored as 18 reals)
    constexpr bool R12T = (MAG_FP64_READ12 & 16) != 0 && std::is_same<TS, double>::value;
    if (!tiled) {
      if (iw)
        r = launch(c, "stout_top", k_stout_top<ACT_IWASAKI, TS, NC, TS, NCN, R12T>,
                   k_stout_top_b<ACT_IWASAKI, TS, NC, TS, NCN, R12T>, grid_stencil(g), tp);
      else
        r = launch(c, "stout_top", k_stout_top<ACT_WILSON, TS, NC, TS, NCN>,
                   k_stout_top_b<ACT_WILSON, TS, NC, TS, NCN>, grid_stencil(g), tp);
    }
    if (r) return r;
  }
  for (int k = n - 1; k >= 1; --k) {
    StoutBwdArgs<TS, NC> bw{};
    bw.g = g;
    bw.U = lview<TS, NC>(Ulev(k), g);
    bw.L = aview<TS>(ws->L[cur], g);
    bw.W = mview<TS>(ws->W, g);
    bw.Ulo = lview<TS, NC>(Ulev(k - 1), g);
    bw.Qlo = aview<TS>(ws->Qs[k - 1], g);
    bw.Clo = mview<TS>(ws->Cs[k - 1], g);
    bw.Lout = aview<TS>(ws->L[1 - cur], g);
    bw.Wout = mview<TS>(ws->W, g);
    bw.rho = act.rho;
    if ((r = launch(c, "stout_bwd", k_stout_bwd<false, TS, NC, TS, R12B>, k_stout_bwd_b<false, TS, NC, TS, R12B>,
                    grid_stencil(g), bw)))
      return r;
    cur = 1 - cur;
  }
  StoutBwdArgs<TS, NC> bw{};
  bw.g = g;
  bw.U = lview<TS, NC>(U, g);
  bw.L = aview<TS>(ws->L[cur], g);
  bw.W = mview<TS>(ws->W, g);
  bw.rho = act.rho;
  bw.f = fa;
  if (act.mag) {
    if ((r = ws->need_Z())) return r;
    bw.Zout = mview<TS>(ws->Z, g);
  }
  if ((r = launch(c, "stout_bwd_force", k_stout_bwd<true, TS, NC, TS, R12L>, k_stout_bwd_b<true, TS, NC, TS, R12L>,
                  grid_stencil(g), bw)))
    return r;
  if (!act.mag) return MAG_OK;
  // constrained: F from the stored Z_0 with the link classes
  void* Fout = (out == OUT_F) ? F : nullptr;
  if (!Fout) {
    if ((r = ws->need_md())) return r;
    Fout = ws->F;
  }
  ForceZArgs<TS, NC> fz{fa, mview<TS>(ws->Z, g)};
  fz.f.F = aview<TS>(Fout, g);
  if ((r = launch(c, "force_from_z", k_force_from_Z<TS, NC, TS>, k_force_from_Z_b<TS, NC, TS>, grid_links(g), fz)))
    return r;
  if (out == OUT_PADD) return mom_update(c, g, std::is_same<TS, double>::value ? MAG_F64 : MAG_F32, P, Fout, eps);
  return MAG_OK;
}

// force_gradient_integrator (field-hmc.h:525-567). The auxiliary field is
// written directly by the first force of each half step (no copy), and the
// second force of each half step updates the momenta in its epilogue.
template <class TS, int NC>
int Ops<TS, NC>::fg(mag_ctx c, const Geom& g, void* U, void* P, const mag_action& act, int steps, double dt,
                    Workspace* ws) {
  int r;
  if ((r = ws->need_md())) return r;
  const double alpha = (3. - std::sqrt(3.)) * dt / 6.;
  const double beta = dt / std::sqrt(3.);
  const double gamma = (2. - std::sqrt(3.)) * dt * dt / 12.;
  if ((r = link_update(c, g, U, U, P, alpha, act.mag))) return r;
  for (int i = 0; i < steps; ++i) {
    for (int half = 0; half < 2; ++half) {
      if ((r = force(c, g, U, act, OUT_UAUX, nullptr, nullptr, ws->Uaux, gamma, ws))) return r;
      if ((r = force(c, g, ws->Uaux, act, OUT_PADD, nullptr, P, nullptr, 0.5 * dt, ws))) return r;
      if (half == 0 && (r = link_update(c, g, U, U, P, beta, act.mag))) return r;
    }
    const double a = (i < steps - 1) ? 2. * alpha : alpha;
    if ((r = link_update(c, g, U, U, P, a, act.mag))) return r;
  }
  return MAG_OK;
}

// leap_frog_integrator (field-hmc.h:569-587).
template <class TS, int NC>
int Ops<TS, NC>::leapfrog(mag_ctx c, const Geom& g, void* U, void* P, const mag_action& act, int steps, double dt,
                          Workspace* ws) {
  int r;
  if ((r = ws->need_md())) return r;
  if ((r = link_update(c, g, U, U, P, dt / 2., act.mag))) return r;
  for (int i = 0; i < steps; ++i) {
    if ((r = force(c, g, U, act, OUT_PADD, nullptr, P, nullptr, dt, ws))) return r;
    const double a = (i < steps - 1) ? dt : dt / 2.;
    if ((r = link_update(c, g, U, U, P, a, act.mag))) return r;
  }
  return MAG_OK;
}

template <class TS, int NC>
int Ops<TS, NC>::action_sums(mag_ctx c, const Geom& g, void* U, bool rect, int slot) {
  const int nb = (g.V + BLOCK - 1) / BLOCK;
  int r;
  if ((r = ensure_part(c, 2 * (size_t)nb))) return r;
  ActionArgs<TS, NC> aa{lview<TS, NC>(U, g), g, c->d_part, rect ? 1 : 0};
  if ((r = launch(c, "action", k_action_partial<TS, NC>, k_action_partial_b<TS, NC>, dim3(nb), aa))) return r;
  if ((r = final_reduce(c, c->d_part, nb, 2, 0, slot))) return r;
  return final_reduce(c, c->d_part + 1, nb, 2, 0, slot + 1);
}

// The Hamiltonian's action on the smeared field: smearing in FP64 from the
// stored links (FP32 or FP64), then the FP64 plaquette/rectangle sums.
template <class TS, int NC>
int Ops<TS, NC>::smeared_action_sums(mag_ctx c, const Geom& g, void* U, double rho, int n, bool rect, int slot) {
  if (n <= 0) return action_sums(c, g, U, rect, slot);
  Workspace* wsd = workspace(c, g, MAG_F64, 18);
  int r;
  if ((r = wsd->need_h(n))) return r;
  for (int k = 0; k < n; ++k) {
    void* dst = wsd->Uh[k & 1];
    if (k == 0) {
      StoutFwdArgs<TS, NC, double, 18> fw{g, lview<TS, NC>(U, g), aview<double>(nullptr, g),
                                          mview<double>(nullptr, g), lview<double, 18>(dst, g), rho, 0};
      r = launch(c, "stout_fwd_h", k_stout_fwd<TS, NC, double, 18, double>, k_stout_fwd_b<TS, NC, double, 18, double>,
                 grid_stencil(g), fw);
    } else {
      StoutFwdArgs<double, 18, double, 18> fw{g, lview<double, 18>(wsd->Uh[(k - 1) & 1], g),
                                              aview<double>(nullptr, g), mview<double>(nullptr, g),
                                              lview<double, 18>(dst, g), rho, 0};
      r = launch(c, "stout_fwd_h", k_stout_fwd<double, 18, double, 18, double>,
                 k_stout_fwd_b<double, 18, double, 18, double>, grid_stencil(g), fw);
    }
    if (r) return r;
  }
  void* top = wsd->Uh[(n - 1) & 1];
  const int nb = (g.V + BLOCK - 1) / BLOCK;
  if ((r = ensure_part(c, 2 * (size_t)nb))) return r;
  ActionArgs<double, 18> aa{lview<double, 18>(top, g), g, c->d_part, rect ? 1 : 0};
  if ((r = launch(c, "action", k_action_partial<double, 18>, k_action_partial_b<double, 18>, dim3(nb), aa))) return r;
  if ((r = final_reduce(c, c->d_part, nb, 2, 0, slot))) return r;
  return final_reduce(c, c->d_part + 1, nb, 2, 0, slot + 1);
}

template <class TS, int NC>
int Ops<TS, NC>::smear(mag_ctx c, const Geom& g, void* Uout, void* U, double rho, int n, Workspace* ws) {
  int r;
  if ((r = ws->need_md())) return r;
  void* src = U;
  for (int k = 0; k < n; ++k) {
    void* dst = (k == n - 1) ? Uout : ((src == ws->Uaux) ? Uout : ws->Uaux);
    if (dst == src) dst = (src == Uout) ? ws->Uaux : Uout;
    StoutFwdArgs<TS, NC, TS, NC> fw{g, lview<TS, NC>(src, g), aview<TS>(nullptr, g), mview<TS>(nullptr, g),
                                    lview<TS, NC>(dst, g), rho, 0};
    if ((r = launch(c, "stout_fwd", k_stout_fwd<TS, NC, TS, NC, TS>, k_stout_fwd_b<TS, NC, TS, NC, TS>, grid_stencil(g),
                    fw)))
      return r;
    src = dst;
  }
  if (src != Uout) return copy_dev(c, Uout, src, (size_t)g.V * 4 * NC * sizeof(TS));
  return MAG_OK;
}

template <class TS, int NC>
int Ops<TS, NC>::reunitarize(mag_ctx c, const Geom& g, void* U, int slot) {
  const int nb = (g.V + BLOCK - 1) / BLOCK;
  int r;
  if ((r = ensure_part(c, 4 * (size_t)nb))) return r;
  ReunitArgs<TS, NC> ra{lview<TS, NC>(U, g), g.V, c->d_part};
  if ((r = launch(c, "reunitarize", k_reunitarize<TS, NC>, k_reunitarize_b<TS, NC>, grid_links(g), ra))) return r;
  return final_reduce(c, c->d_part, 4 * nb, 1, 1, slot);
}

template <class TS, int NC>
int Ops<TS, NC>::max_diff(mag_ctx c, const Geom& g, void* A, void* B, int slot) {
  const int nb = (g.V + BLOCK - 1) / BLOCK;
  int r;
  if ((r = ensure_part(c, 4 * (size_t)nb))) return r;
  hipLaunchKernelGGL((k_max_diff<TS, NC>), grid_links(g), dim3(BLOCK), 0, c->stream, lview<TS, NC>(A, g),
                     lview<TS, NC>(B, g), g.V, c->d_part);
  CHECK_LAUNCH(c);
  return final_reduce(c, c->d_part, 4 * nb, 1, 1, slot);
}

template <class TS, int NC>
int Ops<TS, NC>::haar(mag_ctx c, const Geom& g, void* U, uint64_t seed) {
  hipLaunchKernelGGL((k_haar<TS, NC>), grid_links(g), dim3(BLOCK), 0, c->stream, lview<TS, NC>(U, g), g.V, seed);
  CHECK_LAUNCH(c);
  return MAG_OK;
}

template <class TS, int NC>
int Ops<TS, NC>::unit(mag_ctx c, const Geom& g, void* U) {
  hipLaunchKernelGGL((k_unit<TS, NC>), grid_links(g), dim3(BLOCK), 0, c->stream, lview<TS, NC>(U, g), g.V);
  CHECK_LAUNCH(c);
  return MAG_OK;
}

template <class TS, int NC>
int Ops<TS, NC>::wilson_loop(mag_ctx c, const Geom& g, void* U, const int r3[3], int T, int slot, int* nperm) {
  WloopArgs a{};
  a.g = g;
  a.T = T;
  int rl[3] = {r3[0], r3[1], r3[2]};
  std::sort(rl, rl + 3);
  a.nperm = 0;
  do {
    for (int i = 0; i < 3; ++i) a.r[a.nperm][i] = rl[i];
    a.nperm++;
  } while (std::next_permutation(rl, rl + 3));
  const int nb = (g.V + BLOCK - 1) / BLOCK;
  int r;
  if ((r = ensure_part(c, (size_t)nb))) return r;
  {
    KTimer t(c, "wilson_loop");
    hipLaunchKernelGGL((k_wilson_loop<TS, NC>), dim3(nb), dim3(BLOCK), 0, c->stream, lview<TS, NC>(U, g), a, c->d_part);
  }
  CHECK_LAUNCH(c);
  *nperm = a.nperm;
  return final_reduce(c, c->d_part, nb, 1, 0, slot);
}

}  // namespace mag
