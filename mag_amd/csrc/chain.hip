// Device-resident HMC chain: run_chmc / update_field (field-hmc.h:770-986).
// Links, backup, momenta and workspace stay in HBM; each trajectory issues
// its kernels on the context stream and synchronises once, to read the eight
// FP64 scalars (H_old parts, H_new parts, plaquettes, reunitarization) that
// the accept step needs.
#include <cmath>
#include <cstring>

#include "runtime.hpp"

using namespace mag;

struct mag_chain_s {
  mag_ctx ctx = nullptr;
  int dims[4] = {0, 0, 0, 0};
  Geom g;
  mag_action act{};
  mag_hmc_params p{};
  int prec = MAG_F64, kind = MAG_LINK18;
  mag_field U = nullptr, Ubak = nullptr, P = nullptr;
  // FP64 fields of the reversibility check (created on first use)
  mag_field U64 = nullptr, U64b = nullptr, P64 = nullptr;
  long traj = 0;
  bool demon_init = false;
  double demon[4] = {0., 0., 0., 0.};  // S0p, S0r, Ep, Er
  double emax = 0.;
  int naccept = 0;
};

namespace {

enum : uint32_t { STREAM_ACCEPT = 3 };

// scalar slots of one trajectory
enum { S_OLD_P = 0, S_OLD_R = 1, S_OLD_K = 2, S_NEW_P = 4, S_NEW_R = 5, S_NEW_K = 6, S_DEV = 7, S_THIN_OLD = 8,
       S_THIN_NEW = 10 };

template <class TS, int NC>
int traj_kernels(mag_chain c, bool need_rect) {
  using O = Ops<TS, NC>;
  mag_ctx x = c->ctx;
  const Geom& g = c->g;
  Workspace* ws = workspace(x, g, c->prec, NC);
  const uint64_t i = (uint64_t)c->traj;
  int r;
  // init_momentum (field-hmc.h:810)
  if ((r = mom_refresh(x, g, c->prec, c->P->p, c->p.seed, i, 1.))) return r;
  // H_old (field-hmc.h:812) and the thin plaquette of the old state
  if ((r = O::smeared_action_sums(x, g, c->U->p, c->act.rho, c->act.n_stout, need_rect, S_OLD_P))) return r;
  if ((r = kinetic_sum(x, g, c->prec, c->P->p, c->act.mag, S_OLD_K))) return r;
  if ((r = O::action_sums(x, g, c->U->p, false, S_THIN_OLD))) return r;
  // backup, integrate (field-hmc.h:814)
  if ((r = copy_dev(x, c->Ubak->p, c->U->p, c->U->bytes))) return r;
  if (c->p.integrator == 0) r = O::fg(x, g, c->U->p, c->P->p, c->act, c->p.trajectory_length, c->p.dt, ws);
  else r = O::leapfrog(x, g, c->U->p, c->P->p, c->act, c->p.trajectory_length, c->p.dt, ws);
  if (r) return r;
  if (c->p.reunitarize) {
    if ((r = O::reunitarize(x, g, c->U->p, S_DEV))) return r;
  } else {
    if ((r = fill_dev(x, x->d_scalar + S_DEV, 0., 1))) return r;
  }
  // H_new (field-hmc.h:816)
  if ((r = O::smeared_action_sums(x, g, c->U->p, c->act.rho, c->act.n_stout, need_rect, S_NEW_P))) return r;
  if ((r = kinetic_sum(x, g, c->prec, c->P->p, c->act.mag, S_NEW_K))) return r;
  return O::action_sums(x, g, c->U->p, false, S_THIN_NEW);
}

int run_one(mag_chain c, mag_traj_result* out) {
  mag_ctx x = c->ctx;
  const bool need_rect = c->act.type == MAG_IWASAKI || c->p.accept_mode != 0;
  int r = dispatch_link(x, c->prec, c->kind, [&](auto t) -> int {
    return traj_kernels<typename decltype(t)::T, decltype(t)::N>(c, need_rect);
  });
  if (r) return r;
  if ((r = fetch_scalars(x, 12))) return r;
  const double* s = x->h_scalar;
  const double S_old = action_from_sums(c->act, s[S_OLD_P], s[S_OLD_R]);
  const double S_new = action_from_sums(c->act, s[S_NEW_P], s[S_NEW_R]);
  const double h_old = s[S_OLD_K] / 2. + S_old;
  const double h_new = s[S_NEW_K] / 2. + S_new;
  const long i = c->traj;
  // accept step (field-hmc.h:818-843)
  const double die = host_uniform(c->p.seed, STREAM_ACCEPT, (uint64_t)i);
  const double prob = std::exp(h_old - h_new);
  bool acc = die < prob;
  const double Sp = -s[S_NEW_P], Sr = -s[S_NEW_R];
  if (c->p.accept_mode != 0) {
    // demon window on the raw plaquette / rectangle sums (field-demon.h:173-176)
    const double e = c->emax;
    const bool win = (c->demon[0] - Sp < e) && (c->demon[0] - Sp > -e) && (c->demon[1] - Sr < e) &&
                     (c->demon[1] - Sr > -e);
    acc = (c->p.accept_mode == 1) ? win : (acc && win);
  }
  const bool bad = std::isnan(h_new) || std::isnan(h_old);
  if (i < c->p.num_forced_accept && !bad) acc = true;
  if (bad) acc = false;
  if (acc) {
    if (c->p.accept_mode != 0) {
      c->demon[2] = c->demon[0] - Sp;
      c->demon[3] = c->demon[1] - Sr;
    }
    c->naccept++;
  } else {
    HIPCHK(x, hipMemcpyAsync(c->U->p, c->Ubak->p, c->U->bytes, hipMemcpyDeviceToDevice, x->stream));
  }
  if (out) {
    out->h_old = h_old;
    out->h_new = h_new;
    out->dh = h_new - h_old;
    out->die_roll = die;
    out->accept = acc ? 1 : 0;
    out->avg_plaq = (acc ? s[S_THIN_NEW] : s[S_THIN_OLD]) / (18. * c->g.V);
    out->s_gauge = acc ? S_new : S_old;
    out->demon_e[0] = c->demon[2];
    out->demon_e[1] = c->demon[3];
    out->reunit_dev = s[S_DEV];
    out->status = bad ? MAG_ERR_NAN : MAG_OK;
  }
  c->traj++;
  return MAG_OK;
}

int init_demon(mag_chain c) {
  if (c->demon_init) return MAG_OK;
  mag_ctx x = c->ctx;
  int r = dispatch_link(x, c->prec, c->kind, [&](auto t) -> int {
    using O = Ops<typename decltype(t)::T, decltype(t)::N>;
    return O::smeared_action_sums(x, c->g, c->U->p, c->act.rho, c->act.n_stout, true, 0);
  });
  if (r || (r = fetch_scalars(x, 2))) return r;
  c->demon[0] = -x->h_scalar[0];
  c->demon[1] = -x->h_scalar[1];
  c->demon[2] = c->demon[3] = 0.;
  c->emax = c->p.demon_emax >= 0. ? c->p.demon_emax : -c->p.demon_emax * std::fabs(c->demon[0]);
  c->demon_init = true;
  return MAG_OK;
}

}  // namespace

extern "C" {

int mag_chain_create(mag_ctx ctx, const int dims[4], const mag_action* act, const mag_hmc_params* p,
                     mag_chain* out) {
  if (!ctx || !dims || !act || !p || !out) return MAG_ERR_ARG;
  *out = nullptr;
  if (p->trajectory_length < 0 || !(p->dt >= 0.) || (p->integrator != 0 && p->integrator != 1) ||
      p->accept_mode < 0 || p->accept_mode > 2 || act->n_stout < 0 || act->mag < 0)
    return fail(ctx, MAG_ERR_ARG, "bad chain parameters");
  auto* c = new mag_chain_s();
  c->ctx = ctx;
  std::memcpy(c->dims, dims, sizeof c->dims);
  c->g = make_geom(dims);
  c->act = *act;
  c->p = *p;
  c->prec = p->precision == MAG_F32 ? MAG_F32 : MAG_F64;
  c->kind = c->prec == MAG_F32 ? MAG_LINK12 : MAG_LINK18;
  int r;
  if ((r = mag_field_create(ctx, dims, c->kind, c->prec, &c->U)) ||
      (r = mag_field_create(ctx, dims, c->kind, c->prec, &c->Ubak)) ||
      (r = mag_field_create(ctx, dims, MAG_ALG8, c->prec, &c->P)) || (r = mag_unit_links(c->U))) {
    mag_chain_destroy(c);
    return r;
  }
  *out = c;
  return MAG_OK;
}

int mag_chain_destroy(mag_chain c) {
  if (!c) return MAG_ERR_ARG;
  for (mag_field f : {c->U, c->Ubak, c->P, c->U64, c->U64b, c->P64})
    if (f) mag_field_destroy(f);
  delete c;
  return MAG_OK;
}

int mag_chain_set_links_aos(mag_chain c, const double* host) {
  if (!c || !host) return MAG_ERR_ARG;
  c->demon_init = false;
  return mag_field_upload_aos(c->U, host);
}

int mag_chain_get_links_aos(mag_chain c, double* host) {
  if (!c || !host) return MAG_ERR_ARG;
  return mag_field_download_aos(c->U, host);
}

int mag_chain_haar_links(mag_chain c, uint64_t seed) {
  if (!c) return MAG_ERR_ARG;
  c->demon_init = false;
  return mag_haar_links(c->U, seed);
}

int mag_chain_links(mag_chain c, mag_field* U) {
  if (!c || !U) return MAG_ERR_ARG;
  *U = c->U;
  return MAG_OK;
}

int mag_chain_run(mag_chain c, int ntraj, mag_traj_result* results) {
  if (!c || ntraj < 0) return MAG_ERR_ARG;
  int r;
  if (c->p.accept_mode != 0 && (r = init_demon(c))) return r;
  for (int t = 0; t < ntraj; ++t)
    if ((r = run_one(c, results ? results + t : nullptr))) return r;
  return MAG_OK;
}

int mag_chain_demon_state(mag_chain c, double state[4]) {
  if (!c || !state) return MAG_ERR_ARG;
  int r;
  if ((r = init_demon(c))) return r;
  std::memcpy(state, c->demon, sizeof c->demon);
  return MAG_OK;
}

int mag_chain_traj_index(mag_chain c, long* next) {
  if (!c || !next) return MAG_ERR_ARG;
  *next = c->traj;
  return MAG_OK;
}

// FP64 round trip from the current state: momenta from a stream epoch that no
// trajectory uses (2^40 + trajectory index), FG forward, momenta flipped, FG
// forward again, momenta flipped back. The chain's own fields are untouched.
int mag_chain_reversibility(mag_chain c, double* max_link_dev, double* dh_roundtrip) {
  if (!c) return MAG_ERR_ARG;
  mag_ctx x = c->ctx;
  int r;
  if (!c->U64 && ((r = mag_field_create(x, c->dims, MAG_LINK18, MAG_F64, &c->U64)) ||
                  (r = mag_field_create(x, c->dims, MAG_LINK18, MAG_F64, &c->U64b)) ||
                  (r = mag_field_create(x, c->dims, MAG_ALG8, MAG_F64, &c->P64))))
    return r;
  if ((r = mag_field_copy(c->U64, c->U)) || (r = mag_field_copy(c->U64b, c->U64))) return r;
  const uint64_t epoch = (1ull << 40) + (uint64_t)c->traj;
  if ((r = mom_refresh(x, c->g, MAG_F64, c->P64->p, c->p.seed, epoch, 1.))) return r;
  double parts[2], h0 = 0., h1 = 0.;
  if ((r = mag_hamiltonian(c->U64, c->P64, &c->act, parts, &h0))) return r;
  using O = Ops<double, 18>;
  Workspace* ws = workspace(x, c->g, MAG_F64, 18);
  const int n = c->p.trajectory_length;
  const double dt = c->p.dt;
  if (c->p.integrator == 0) r = O::fg(x, c->g, c->U64->p, c->P64->p, c->act, n, dt, ws);
  else r = O::leapfrog(x, c->g, c->U64->p, c->P64->p, c->act, n, dt, ws);
  if (r || (r = alg_scale(x, c->g, MAG_F64, c->P64->p, -1.))) return r;
  if (c->p.integrator == 0) r = O::fg(x, c->g, c->U64->p, c->P64->p, c->act, n, dt, ws);
  else r = O::leapfrog(x, c->g, c->U64->p, c->P64->p, c->act, n, dt, ws);
  if (r || (r = alg_scale(x, c->g, MAG_F64, c->P64->p, -1.))) return r;
  if ((r = mag_hamiltonian(c->U64, c->P64, &c->act, parts, &h1))) return r;
  if ((r = O::max_diff(x, c->g, c->U64->p, c->U64b->p, 0)) || (r = fetch_scalars(x, 1))) return r;
  if (max_link_dev) *max_link_dev = x->h_scalar[0];
  if (dh_roundtrip) *dh_roundtrip = h1 - h0;
  return MAG_OK;
}

}  // extern "C"
