// Device-resident HMC chain: run_chmc / update_field (field-hmc.h:770-986).
// Links, backup, momenta and workspace stay in HBM; each trajectory issues
// its kernels on the chain's stream and synchronises once, to read the
// twelve FP64 scalars (H_old and H_new parts, thin plaquettes,
// reunitarization deviation) that the accept step needs.
#include <cmath>
#include <cstring>

#include "chain_internal.hpp"

using namespace mag;

namespace mag {

enum : uint32_t { STREAM_ACCEPT = 3 };

bool chain_needs_rect(mag_chain c) { return c->act.type == MAG_IWASAKI || c->p.accept_mode != 0; }

namespace {

template <class TS, int NC>
int traj_kernels(mag_chain c, bool use_epoch_ptr) {
  using O = Ops<TS, NC>;
  mag_ctx x = c->ctx;
  const Geom& g = c->g;
  Workspace* ws = workspace(x, g, c->prec, NC);
  const bool rect = chain_needs_rect(c);
  int r;
  // init_momentum (field-hmc.h:810)
  if ((r = mom_refresh(x, g, c->prec, c->P->p, c->p.seed, (uint64_t)c->traj, 1.,
                       use_epoch_ptr ? c->d_epoch : nullptr)))
    return r;
  // H_old (field-hmc.h:812) and the thin plaquette of the old state; the
  // action sums come from the previous accept step when it left them (not
  // while a serving batch records its fixed launch sequence)
  c->s_used = c->s_valid && !x->rec;
  if (!c->s_used) {
    if ((r = O::smeared_action_sums(x, g, c->U->p, c->act.rho, c->act.n_stout, rect, S_OLD_P))) return r;
    if ((r = O::action_sums(x, g, c->U->p, false, S_THIN_OLD))) return r;
  }
  if ((r = kinetic_sum(x, g, c->prec, c->P->p, c->act.mag, S_OLD_K))) return r;
  // backup, integrate (field-hmc.h:814)
  if ((r = copy_dev(x, c->Ubak->p, c->U->p, c->U->bytes))) return r;
  if (c->p.integrator == 0) r = O::fg(x, g, c->U->p, c->P->p, c->act, c->p.trajectory_length, c->p.dt, ws);
  else r = O::leapfrog(x, g, c->U->p, c->P->p, c->act, c->p.trajectory_length, c->p.dt, ws);
  if (r) return r;
  if (c->p.reunitarize) r = O::reunitarize(x, g, c->U->p, S_DEV);
  else r = fill_dev(x, x->d_scalar + S_DEV, 0., 1);
  if (r) return r;
  // H_new (field-hmc.h:816)
  if ((r = O::smeared_action_sums(x, g, c->U->p, c->act.rho, c->act.n_stout, rect, S_NEW_P))) return r;
  if ((r = kinetic_sum(x, g, c->prec, c->P->p, c->act.mag, S_NEW_K))) return r;
  return O::action_sums(x, g, c->U->p, false, S_THIN_NEW);
}

}  // namespace

int chain_traj_kernels(mag_chain c, bool use_epoch_ptr) {
  return dispatch_link(c->ctx, c->prec, c->kind, [&](auto t) -> int {
    return traj_kernels<typename decltype(t)::T, decltype(t)::N>(c, use_epoch_ptr);
  });
}

int chain_accept(mag_chain c, const double* s_in, mag_traj_result* out) {
  mag_ctx x = c->ctx;
  double s[S_COUNT];
  std::memcpy(s, s_in, sizeof s);
  if (c->s_used) {
    s[S_OLD_P] = c->s_cache[0];
    s[S_OLD_R] = c->s_cache[1];
    s[S_THIN_OLD] = c->s_cache[2];
  }
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
    const bool win =
        (c->demon[0] - Sp < e) && (c->demon[0] - Sp > -e) && (c->demon[1] - Sr < e) && (c->demon[1] - Sr > -e);
    acc = (c->p.accept_mode == 1) ? win : (acc && win);
  }
  // a non-finite Hamiltonian is rejected, also during forced accepts
  const bool bad = !std::isfinite(h_new) || !std::isfinite(h_old);
  if (i < c->p.num_forced_accept && !bad) acc = true;
  if (bad) acc = false;
  if (acc) {
    if (c->p.accept_mode != 0) {
      c->demon[2] = c->demon[0] - Sp;
      c->demon[3] = c->demon[1] - Sr;
    }
    c->naccept++;
  } else {
    int r = copy_dev(x, c->U->p, c->Ubak->p, c->U->bytes);
    if (r) return r;
  }
  // the action sums of the links the chain now holds
  const int k0 = acc ? S_NEW_P : S_OLD_P, k1 = acc ? S_NEW_R : S_OLD_R, k2 = acc ? S_THIN_NEW : S_THIN_OLD;
  c->s_cache[0] = s[k0];
  c->s_cache[1] = s[k1];
  c->s_cache[2] = s[k2];
  c->s_valid = std::isfinite(s[k0]) && std::isfinite(s[k1]) && std::isfinite(s[k2]);
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

int chain_init_demon(mag_chain c) {
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

int ctx_child_create(mag_ctx parent, hipStream_t stream, std::shared_ptr<MemPool> pool, mag_ctx* out) {
  auto* c = new mag_ctx_s();
  c->device = parent->device;
  c->stream = stream;
  c->owns_stream = false;
  c->pool = std::move(pool);
  if (dev_alloc(c, (void**)&c->d_scalar, 16 * sizeof(double)) ||
      host_alloc(c, (void**)&c->h_scalar, 16 * sizeof(double))) {
    dev_free(c, c->d_scalar, 16 * sizeof(double));
    parent->err = c->err;
    delete c;
    return MAG_ERR_OOM;
  }
  *out = c;
  return MAG_OK;
}

size_t chain_footprint(const int dims[4], const mag_action& act, const mag_hmc_params& p) {
  const size_t V = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
  const bool f32 = p.precision == MAG_F32;
  const size_t rs = f32 ? 4 : 8, nc = (f32 || p.recon12) ? 12 : 18;
  const size_t link = V * 4 * nc * rs, alg = V * 4 * 8 * rs, mat = V * 4 * 18 * rs, mat64 = V * 4 * 18 * 8;
  size_t b = 3 * link + 2 * alg;  // U, Ubak, Uaux; P, F
  if (act.n_stout > 0) {
    b += (size_t)act.n_stout * (link + alg + mat) + 2 * alg + mat;  // levels; Lambda[2], W
    b += (act.n_stout > 1 ? 2 : 1) * mat64;                    // FP64 smearing for H
    if (!f32 && p.recon12) b += mat64;                         // FP64 12-real: top level as 18 reals
    if (act.mag) b += mat;
  }
  b += ((V + 255) / 256 * 4 + 4096) * sizeof(double) + 64 * sizeof(double);
  return b;
}

}  // namespace mag

extern "C" {

int mag_chain_create(mag_ctx ctx, const int dims[4], const mag_action* act, const mag_hmc_params* p,
                     mag_chain* out) {
  if (!ctx || !dims || !act || !p || !out) return MAG_ERR_ARG;
  *out = nullptr;
  if (p->trajectory_length < 0 || !(p->dt >= 0.) || (p->integrator != 0 && p->integrator != 1) ||
      p->accept_mode < 0 || p->accept_mode > 2 || act->n_stout < 0 || act->mag < 0 ||
      (act->type != MAG_WILSON && act->type != MAG_IWASAKI))
    return fail(ctx, MAG_ERR_ARG, "bad chain parameters");
  auto* c = new mag_chain_s();
  c->ctx = ctx;
  std::memcpy(c->dims, dims, sizeof c->dims);
  c->g = make_geom(dims);
  c->act = *act;
  c->p = *p;
  c->prec = p->precision == MAG_F32 ? MAG_F32 : MAG_F64;
  // FP32 links are always 12-real; FP64 links 12-real when recon12 is set
  c->kind = (c->prec == MAG_F32 || p->recon12) ? MAG_LINK12 : MAG_LINK18;
  int r;
  if ((r = mag_field_create(ctx, dims, c->kind, c->prec, &c->U)) ||
      (r = mag_field_create(ctx, dims, c->kind, c->prec, &c->Ubak)) ||
      (r = mag_field_create(ctx, dims, MAG_ALG8, c->prec, &c->P)) ||
      (r = dev_alloc(ctx, (void**)&c->d_epoch, 16)) ||
      // the largest partial-sum need of a trajectory (reunitarize: 4 per block)
      (r = ensure_part(ctx, 4 * (size_t)((c->g.V + RT_BLOCK - 1) / RT_BLOCK))) || (r = mag_unit_links(c->U))) {
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
  if (c->d_epoch) dev_free(c->ctx, c->d_epoch, 16);
  if (c->owns_ctx) mag_ctx_destroy(c->ctx);
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

// Links written through the borrowed handle: drop the demon window centre
// (and demon energies) so the next run or demon_state call takes them from
// the links the chain now holds.
int mag_chain_reset_demon(mag_chain c) {
  if (!c) return MAG_ERR_ARG;
  c->demon_init = false;
  c->s_valid = false;
  return MAG_OK;
}

int mag_chain_run(mag_chain c, int ntraj, mag_traj_result* results) {
  if (!c || ntraj < 0) return MAG_ERR_ARG;
  int r;
  if (c->p.accept_mode != 0 && (r = chain_init_demon(c))) return r;
  c->s_valid = false;  // the links may have been written since the last call
  for (int t = 0; t < ntraj; ++t) {
    if ((r = chain_traj_kernels(c, false))) return r;
    if ((r = fetch_scalars(c->ctx, S_COUNT))) return r;
    if ((r = chain_accept(c, c->ctx->h_scalar, results ? results + t : nullptr))) return r;
  }
  return MAG_OK;
}

int mag_chain_demon_state(mag_chain c, double state[4]) {
  if (!c || !state) return MAG_ERR_ARG;
  int r;
  if ((r = chain_init_demon(c))) return r;
  std::memcpy(state, c->demon, sizeof c->demon);
  return MAG_OK;
}

int mag_chain_traj_index(mag_chain c, long* next) {
  if (!c || !next) return MAG_ERR_ARG;
  *next = c->traj;
  return MAG_OK;
}

// FP64 round trip from the current state: the links widened to FP64 and
// projected onto SU(3) (FP32 links are off the FP64 group by ~1e-7, and the
// FP64 link update reads links as two rows, which would project them in its
// first step and not in the step back), momenta from a stream epoch that no
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
  using O = Ops<double, 18>;
  if ((r = mag_field_copy(c->U64, c->U)) || (r = O::reunitarize(x, c->g, c->U64->p, 0)) ||
      (r = mag_field_copy(c->U64b, c->U64)))
    return r;
  const uint64_t epoch = (1ull << 40) + (uint64_t)c->traj;
  if ((r = mom_refresh(x, c->g, MAG_F64, c->P64->p, c->p.seed, epoch, 1.))) return r;
  double parts[2], h0 = 0., h1 = 0.;
  if ((r = mag_hamiltonian(c->U64, c->P64, &c->act, parts, &h0))) return r;
  Workspace* ws = workspace(x, c->g, MAG_F64, 18);
  const int n = c->p.trajectory_length;
  const double dt = c->p.dt;
  for (int pass = 0; pass < 2; ++pass) {
    if (c->p.integrator == 0) r = O::fg(x, c->g, c->U64->p, c->P64->p, c->act, n, dt, ws);
    else r = O::leapfrog(x, c->g, c->U64->p, c->P64->p, c->act, n, dt, ws);
    if (r || (r = alg_scale(x, c->g, MAG_F64, c->P64->p, -1.))) return r;
  }
  if ((r = mag_hamiltonian(c->U64, c->P64, &c->act, parts, &h1))) return r;
  if ((r = O::max_diff(x, c->g, c->U64->p, c->U64b->p, 0)) || (r = fetch_scalars(x, 1))) return r;
  if (max_link_dev) *max_link_dev = x->h_scalar[0];
  if (dh_roundtrip) *dh_roundtrip = h1 - h0;
  return MAG_OK;
}

}  // extern "C"
