// C ABI of libmag_amd.so (include/mag_amd.h): contexts, device fields and the
// field-level operations. Arithmetic runs in the gfx950 kernels behind
// Ops<TS, NC> and the common ops; this file validates arguments, picks the
// storage format and moves scalars back to the host.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "kernels.hpp"
#include "runtime.hpp"
#include "su3.hpp"

using namespace mag;

namespace {

bool same_geom(mag_field a, mag_field b) {
  for (int i = 0; i < 4; ++i)
    if (a->dims[i] != b->dims[i]) return false;
  return true;
}

int check_dims(mag_ctx c, const int* d) {
  long v = 1;
  for (int i = 0; i < 4; ++i) {
    if (d[i] < 4) return fail(c, MAG_ERR_GEOMETRY, "lattice extents must be >= 4 (rectangles span 2 links)");
    v *= d[i];
  }
  if (v * 4 * 18 > (1L << 31) - 1 && v * 4 > (1L << 31) - 1)
    return fail(c, MAG_ERR_GEOMETRY, "lattice too large for 32-bit link indices");
  return MAG_OK;
}

bool is_link(mag_field f) { return f && (f->kind == MAG_LINK18 || f->kind == MAG_LINK12); }
bool is_alg(mag_field f) { return f && f->kind == MAG_ALG8; }

template <class TS>
void aos_to_soa(const double* h, TS* s, const Geom& g, int kind) {
  const int nc = ncomp(kind);
  if (kind == MAG_REAL1) {
    for (int site = 0; site < g.V; ++site)
      for (int mu = 0; mu < 4; ++mu) s[(size_t)mu * g.V + site] = TS(h[(size_t)site * 4 + mu]);
    return;
  }
  const int w = pack_width_rt((int)sizeof(TS), nc);
  for (int site = 0; site < g.V; ++site)
    for (int mu = 0; mu < 4; ++mu) {
      const double* m = h + ((size_t)site * 4 + mu) * 18;
      if (kind == MAG_ALG8) {
        M3<double> M;
        for (int k = 0; k < 9; ++k) M.e[k] = {m[2 * k], m[2 * k + 1]};
        double a[8];
        mat_to_alg(M, a);
        for (int k = 0; k < 8; ++k) s[elem_index(mu, k, site, 8, w, g.V)] = TS(a[k]);
      } else {
        for (int k = 0; k < nc; ++k) s[elem_index(mu, k, site, nc, w, g.V)] = TS(m[k]);
      }
    }
}

template <class TS>
void soa_to_aos(const TS* s, double* h, const Geom& g, int kind) {
  const int nc = ncomp(kind);
  if (kind == MAG_REAL1) {
    for (int site = 0; site < g.V; ++site)
      for (int mu = 0; mu < 4; ++mu) h[(size_t)site * 4 + mu] = double(s[(size_t)mu * g.V + site]);
    return;
  }
  const int w = pack_width_rt((int)sizeof(TS), nc);
  for (int site = 0; site < g.V; ++site)
    for (int mu = 0; mu < 4; ++mu) {
      double* m = h + ((size_t)site * 4 + mu) * 18;
      M3<double> M;
      if (kind == MAG_ALG8) {
        double a[8];
        for (int k = 0; k < 8; ++k) a[k] = double(s[elem_index(mu, k, site, 8, w, g.V)]);
        M = alg_to_mat(a);
      } else {
        for (int k = 0; k < nc / 2; ++k)
          M.e[k] = {double(s[elem_index(mu, 2 * k, site, nc, w, g.V)]),
                    double(s[elem_index(mu, 2 * k + 1, site, nc, w, g.V)])};
        if (nc == 12) reconstruct_row2(M);
      }
      for (int k = 0; k < 9; ++k) {
        m[2 * k] = M.e[k].r;
        m[2 * k + 1] = M.e[k].i;
      }
    }
}

// Device batch for the per-matrix test hooks.
int batch_hook(mag_ctx c, long n, int prec, const std::vector<const double*>& ins, double* out, int which) {
  if (!c || n < 0 || !out) return MAG_ERR_ARG;
  if (n == 0) return MAG_OK;
  const size_t b = (size_t)n * 18 * sizeof(double);
  std::vector<void*> d(ins.size() + 1, nullptr);
  int r = MAG_OK;
  for (size_t i = 0; i < d.size() && r == MAG_OK; ++i) r = dev_alloc(c, &d[i], b);
  if (r == MAG_OK) {
    for (size_t i = 0; i < ins.size() && r == MAG_OK; ++i)
      if (hipMemcpyAsync(d[i], ins[i], b, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        r = fail(c, MAG_ERR_DEVICE, "hook upload failed");
    if (r == MAG_OK) {
      double* o = (double*)d.back();
      r = which == 0   ? batch_expiQ(c, (const double*)d[0], o, n, prec)
          : which == 1 ? batch_lambda(c, (const double*)d[0], (const double*)d[1], (const double*)d[2], o, n, prec)
                       : batch_dexpiq(c, (const double*)d[0], (const double*)d[1], o, n, prec);
      if (r == MAG_OK && (hipMemcpyAsync(out, o, b, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                          hipStreamSynchronize(c->stream) != hipSuccess))
        r = fail(c, MAG_ERR_DEVICE, "hook download failed");
    }
  }
  (void)hipStreamSynchronize(c->stream);
  for (size_t i = 0; i < d.size(); ++i) dev_free(c, d[i], d[i] ? b : 0);
  return r;
}

}  // namespace

extern "C" {

int mag_ctx_create(int device, mag_ctx* out) {
  if (!out) return MAG_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MAG_ERR_DEVICE;
  if (device < 0 || device >= n) return MAG_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return MAG_ERR_DEVICE;
  auto* c = new mag_ctx_s();
  c->device = device;
  // device address of the allocation-free entry point's scratch (once per device)
  if (noalloc_setup(device, &c->err) != MAG_OK) {
    delete c;
    return MAG_ERR_DEVICE;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MAG_ERR_DEVICE;
  }
  if (hipMalloc(&c->d_scalar, 16 * sizeof(double)) != hipSuccess ||
      hipHostMalloc(&c->h_scalar, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MAG_ERR_OOM;
  }
  c->bytes = 16 * sizeof(double);
  *out = c;
  return MAG_OK;
}

int mag_ctx_destroy(mag_ctx c) {
  if (!c) return MAG_ERR_ARG;
  (void)hipSetDevice(c->device);
  // a context on a borrowed stream (serving) is destroyed by its owner
  // once its own work is known to be complete
  if (c->owns_stream) (void)hipStreamSynchronize(c->stream);
  collect_timers(c);
  c->ws.clear();
  dev_free(c, c->d_part, c->part_cap * sizeof(double));
  dev_free(c, c->d_scalar, 16 * sizeof(double));
  host_free(c, c->h_scalar, 16 * sizeof(double));
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  if (c->owns_stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MAG_OK;
}

const char* mag_last_error(mag_ctx c) { return c ? c->err.c_str() : "null context"; }

int mag_ctx_synchronize(mag_ctx c) {
  if (!c) return MAG_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAG_OK;
}

int mag_ctx_memory(mag_ctx c, size_t* bytes) {
  if (!c || !bytes) return MAG_ERR_ARG;
  *bytes = c->bytes;
  return MAG_OK;
}

int mag_ctx_set_timing(mag_ctx c, int on) {
  if (!c) return MAG_ERR_ARG;
  c->timing = on != 0;
  return MAG_OK;
}

int mag_ctx_set_timing_family(mag_ctx c, const char* family) {
  if (!c) return MAG_ERR_ARG;
  c->timing_family = family ? family : "";
  return MAG_OK;
}

int mag_ctx_reset_timing(mag_ctx c) {
  if (!c) return MAG_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_timers(c);
  c->timers.clear();
  return MAG_OK;
}

int mag_ctx_kernel_time(mag_ctx c, const char* family, double* ms, long* launches) {
  if (!c || !family) return MAG_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_timers(c);
  auto it = c->timers.find(family);
  if (ms) *ms = it == c->timers.end() ? 0. : it->second.ms;
  if (launches) *launches = it == c->timers.end() ? 0 : it->second.launches;
  return MAG_OK;
}

int mag_field_create(mag_ctx c, const int dims[4], int kind, int prec, mag_field* out) {
  if (!c || !dims || !out) return MAG_ERR_ARG;
  *out = nullptr;
  int r;
  if ((r = check_dims(c, dims))) return r;
  if (kind < MAG_LINK18 || kind > MAG_REAL1 || (prec != MAG_F32 && prec != MAG_F64))
    return fail(c, MAG_ERR_ARG, "bad field kind or precision");
  if (kind == MAG_REAL1 && prec != MAG_F64) return fail(c, MAG_ERR_UNSUPPORTED, "real-per-link fields are FP64");
  if (kind == MAG_LINK18 && prec == MAG_F32)
    return fail(c, MAG_ERR_UNSUPPORTED, "link fields are FP64 18- or 12-real, or FP32 12-real");
  auto* f = new mag_field_s();
  f->ctx = c;
  std::memcpy(f->dims, dims, sizeof f->dims);
  f->kind = kind;
  f->prec = prec;
  f->g = make_geom(dims);
  f->bytes = (size_t)f->g.V * 4 * ncomp(kind) * real_size(prec);
  if ((r = dev_alloc(c, &f->p, f->bytes))) {
    delete f;
    return r;
  }
  if (hipMemsetAsync(f->p, 0, f->bytes, c->stream) != hipSuccess) {
    dev_free(c, f->p, f->bytes);
    delete f;
    return fail(c, MAG_ERR_DEVICE, "field clear failed");
  }
  *out = f;
  return MAG_OK;
}

int mag_field_destroy(mag_field f) {
  if (!f) return MAG_ERR_ARG;
  if (f->owned) {
    // pooled (serving) memory is returned only after its chain's work completed
    if (!f->ctx->pool) (void)hipStreamSynchronize(f->ctx->stream);
    dev_free(f->ctx, f->p, f->bytes);
  }
  delete f;
  return MAG_OK;
}

int mag_field_info(mag_field f, int dims[4], int* kind, int* prec, size_t* bytes) {
  if (!f) return MAG_ERR_ARG;
  if (dims) std::memcpy(dims, f->dims, sizeof f->dims);
  if (kind) *kind = f->kind;
  if (prec) *prec = f->prec;
  if (bytes) *bytes = f->bytes;
  return MAG_OK;
}

int mag_field_upload_aos(mag_field f, const double* host) {
  if (!f || !host) return MAG_ERR_ARG;
  mag_ctx c = f->ctx;
  std::vector<char> stage(f->bytes);
  if (f->prec == MAG_F64) aos_to_soa<double>(host, (double*)stage.data(), f->g, f->kind);
  else aos_to_soa<float>(host, (float*)stage.data(), f->g, f->kind);
  HIPCHK(c, hipMemcpyAsync(f->p, stage.data(), f->bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAG_OK;
}

int mag_field_download_aos(mag_field f, double* host) {
  if (!f || !host) return MAG_ERR_ARG;
  mag_ctx c = f->ctx;
  std::vector<char> stage(f->bytes);
  HIPCHK(c, hipMemcpyAsync(stage.data(), f->p, f->bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (f->prec == MAG_F64) soa_to_aos<double>((const double*)stage.data(), host, f->g, f->kind);
  else soa_to_aos<float>((const float*)stage.data(), host, f->g, f->kind);
  return MAG_OK;
}

int mag_field_copy(mag_field dst, mag_field src) {
  if (!dst || !src || !same_geom(dst, src) || dst->ctx != src->ctx) return MAG_ERR_ARG;
  mag_ctx c = dst->ctx;
  if (dst->kind == src->kind && dst->prec == src->prec) {
    HIPCHK(c, hipMemcpyAsync(dst->p, src->p, dst->bytes, hipMemcpyDeviceToDevice, c->stream));
    return MAG_OK;
  }
  if (!is_link(dst) || !is_link(src)) return fail(c, MAG_ERR_ARG, "only link fields convert between formats");
  return convert_links(c, dst->g, src->prec, src->kind, src->p, dst->prec, dst->kind, dst->p);
}

int mag_unit_links(mag_field U) {
  if (!is_link(U)) return MAG_ERR_ARG;
  return dispatch_link(U->ctx, U->prec, U->kind, [&](auto t) -> int {
    return Ops<typename decltype(t)::T, decltype(t)::N>::unit(U->ctx, U->g, U->p);
  });
}

int mag_haar_links(mag_field U, uint64_t seed) {
  if (!is_link(U)) return MAG_ERR_ARG;
  return dispatch_link(U->ctx, U->prec, U->kind, [&](auto t) -> int {
    return Ops<typename decltype(t)::T, decltype(t)::N>::haar(U->ctx, U->g, U->p, seed);
  });
}

int mag_init_momentum(mag_field P, uint64_t seed, uint64_t traj, double m) {
  if (!is_alg(P) || !(m > 0.)) return MAG_ERR_ARG;
  return mom_refresh(P->ctx, P->g, P->prec, P->p, seed, traj, std::sqrt(m));
}

int mag_reunitarize(mag_field U, double* max_dev) {
  if (!is_link(U)) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  int r = dispatch_link(c, U->prec, U->kind, [&](auto t) -> int {
    return Ops<typename decltype(t)::T, decltype(t)::N>::reunitarize(c, U->g, U->p, 0);
  });
  if (r || (r = fetch_scalars(c, 1))) return r;
  if (max_dev) *max_dev = c->h_scalar[0];
  return MAG_OK;
}

int mag_expiQ_batch(mag_ctx c, const double* Q, double* out, long n, int prec) {
  if (!Q) return MAG_ERR_ARG;
  return batch_hook(c, n, prec, {Q}, out, 0);
}

int mag_compute_Lambda_batch(mag_ctx c, const double* Q, const double* S, const double* U, double* out, long n,
                             int prec) {
  if (!Q || !S || !U) return MAG_ERR_ARG;
  return batch_hook(c, n, prec, {Q, S, U}, out, 1);
}

int mag_DexpiQ_batch(mag_ctx c, const double* Q, const double* DQ, double* out, long n, int prec) {
  if (!Q || !DQ) return MAG_ERR_ARG;
  return batch_hook(c, n, prec, {Q, DQ}, out, 2);
}

int mag_gauge_force(mag_field F, mag_field U, const mag_action* act) {
  if (!is_alg(F) || !is_link(U) || !act || F->prec != U->prec || !same_geom(F, U)) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  return dispatch_link(c, U->prec, U->kind, [&](auto t) -> int {
    using O = Ops<typename decltype(t)::T, decltype(t)::N>;
    Workspace* ws = workspace(c, U->g, U->prec, decltype(t)::N);
    return O::force(c, U->g, U->p, *act, OUT_F, F->p, nullptr, nullptr, 0., ws);
  });
}

int mag_evolve_gauge(mag_field U, mag_field P, double eps, int mag) {
  if (!is_link(U) || !is_alg(P) || U->prec != P->prec || !same_geom(U, P) || mag < 0) return MAG_ERR_ARG;
  return dispatch_link(U->ctx, U->prec, U->kind, [&](auto t) -> int {
    return Ops<typename decltype(t)::T, decltype(t)::N>::link_update(U->ctx, U->g, U->p, U->p, P->p, eps, mag);
  });
}

int mag_evolve_momentum(mag_field P, mag_field F, double eps, double m) {
  if (!is_alg(P) || !is_alg(F) || P->prec != F->prec || !same_geom(P, F)) return MAG_ERR_ARG;
  return mom_update(P->ctx, P->g, P->prec, P->p, F->p, eps * m);
}

int mag_stout_smear(mag_field out, mag_field U, double rho, int nlev) {
  if (!is_link(out) || !is_link(U) || out->kind != U->kind || out->prec != U->prec || !same_geom(out, U) || nlev < 0)
    return MAG_ERR_ARG;
  if (nlev == 0) return mag_field_copy(out, U);
  mag_ctx c = U->ctx;
  return dispatch_link(c, U->prec, U->kind, [&](auto t) -> int {
    using O = Ops<typename decltype(t)::T, decltype(t)::N>;
    Workspace* ws = workspace(c, U->g, U->prec, decltype(t)::N);
    return O::smear(c, U->g, out->p, U->p, rho, nlev, ws);
  });
}

int mag_fg_integrate(mag_field U, mag_field P, const mag_action* act, int steps, double dt, int reunit_flag,
                     double* max_dev) {
  if (!is_link(U) || !is_alg(P) || !act || U->prec != P->prec || !same_geom(U, P) || steps < 0) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  int r = dispatch_link(c, U->prec, U->kind, [&](auto t) -> int {
    using O = Ops<typename decltype(t)::T, decltype(t)::N>;
    Workspace* ws = workspace(c, U->g, U->prec, decltype(t)::N);
    int q = O::fg(c, U->g, U->p, P->p, *act, steps, dt, ws);
    if (q || !reunit_flag) return q;
    return O::reunitarize(c, U->g, U->p, 0);
  });
  if (r) return r;
  if (reunit_flag) {
    if ((r = fetch_scalars(c, 1))) return r;
    if (max_dev) *max_dev = c->h_scalar[0];
  } else if (max_dev) {
    *max_dev = 0.;
  }
  return MAG_OK;
}

int mag_leapfrog_integrate(mag_field U, mag_field P, const mag_action* act, int steps, double dt) {
  if (!is_link(U) || !is_alg(P) || !act || U->prec != P->prec || !same_geom(U, P) || steps < 0) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  return dispatch_link(c, U->prec, U->kind, [&](auto t) -> int {
    using O = Ops<typename decltype(t)::T, decltype(t)::N>;
    Workspace* ws = workspace(c, U->g, U->prec, decltype(t)::N);
    return O::leapfrog(c, U->g, U->p, P->p, *act, steps, dt, ws);
  });
}

static int sums(mag_field U, const mag_action& act, bool rect) {
  return dispatch_link(U->ctx, U->prec, U->kind, [&](auto t) -> int {
    using O = Ops<typename decltype(t)::T, decltype(t)::N>;
    return O::smeared_action_sums(U->ctx, U->g, U->p, act.rho, act.n_stout, rect, 0);
  });
}

int mag_total_plaq(mag_field U, double* out) {
  if (!is_link(U)) return MAG_ERR_ARG;
  mag_action thin{MAG_WILSON, 0., 0., 0, 0, 0.};
  int r = sums(U, thin, false);
  if (r || (r = fetch_scalars(U->ctx, 2))) return r;
  if (out) *out = U->ctx->h_scalar[0];
  return MAG_OK;
}

int mag_total_rectangular(mag_field U, double* out) {
  if (!is_link(U)) return MAG_ERR_ARG;
  mag_action thin{MAG_IWASAKI, 0., 0., 0, 0, 0.};
  int r = sums(U, thin, true);
  if (r || (r = fetch_scalars(U->ctx, 2))) return r;
  if (out) *out = U->ctx->h_scalar[1];
  return MAG_OK;
}

int mag_avg_plaquette(mag_field U, double* out) {
  double t = 0.;
  int r = mag_total_plaq(U, &t);
  if (r) return r;
  if (out) *out = t / (18. * U->g.V);
  return MAG_OK;
}

int mag_gauge_action(mag_field U, const mag_action* act, double* S) {
  if (!is_link(U) || !act) return MAG_ERR_ARG;
  int r = sums(U, *act, act->type == MAG_IWASAKI);
  if (r || (r = fetch_scalars(U->ctx, 2))) return r;
  if (S) *S = action_from_sums(*act, U->ctx->h_scalar[0], U->ctx->h_scalar[1]);
  return MAG_OK;
}

int mag_kinetic_energy(mag_field P, int mag, double* K) {
  if (!is_alg(P)) return MAG_ERR_ARG;
  int r = kinetic_sum(P->ctx, P->g, P->prec, P->p, mag, 2);
  if (r || (r = fetch_scalars(P->ctx, 3))) return r;
  if (K) *K = P->ctx->h_scalar[2] / 2.;
  return MAG_OK;
}

int mag_hamiltonian(mag_field U, mag_field P, const mag_action* act, double parts[2], double* H) {
  if (!is_link(U) || !is_alg(P) || !act || !same_geom(U, P) || U->ctx != P->ctx) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  int r = sums(U, *act, act->type == MAG_IWASAKI);
  if (r || (r = kinetic_sum(c, P->g, P->prec, P->p, act->mag, 2))) return r;
  if ((r = fetch_scalars(c, 3))) return r;
  const double K = c->h_scalar[2] / 2.;
  const double S = action_from_sums(*act, c->h_scalar[0], c->h_scalar[1]);
  if (parts) {
    parts[0] = K;
    parts[1] = S;
  }
  if (H) *H = K + S;
  return MAG_OK;
}

int mag_avg_wilson_loop(mag_field U, const int r3[3], int T, double* out) {
  if (!is_link(U) || !r3 || T <= 0) return MAG_ERR_ARG;
  for (int i = 0; i < 3; ++i)
    if (r3[i] < 0) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  int nperm = 1;
  int r = dispatch_link(c, U->prec, U->kind, [&](auto t) -> int {
    return Ops<typename decltype(t)::T, decltype(t)::N>::wilson_loop(c, U->g, U->p, r3, T, 3, &nperm);
  });
  if (r || (r = fetch_scalars(c, 4))) return r;
  if (out) *out = c->h_scalar[3] / (nperm * (double)U->g.V);
  return MAG_OK;
}

// ------------------------------------------------- asynchronous reductions
// The same sums left on the device: d_out is device memory, written in the
// context's stream order; nothing synchronises.
}  // extern "C"

namespace {
// out[i] = sum_j w[i][j] s[j] over the context's reduction slots, written to
// caller memory in stream order.
int lincomb(mag_ctx c, double* d_out, int nout, const double w[3][4]) {
  if (!d_out) return MAG_ERR_ARG;
  LinCombArgs a{c->d_scalar, d_out, nout, {}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) a.w[i][j] = w[i][j];
  hipLaunchKernelGGL(k_lincomb<0>, dim3(1), dim3(64), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  return MAG_OK;
}
}  // namespace

extern "C" {

int mag_total_plaq_async(mag_field U, double* d_out) {
  if (!is_link(U) || !d_out) return MAG_ERR_ARG;
  mag_action thin{MAG_WILSON, 0., 0., 0, 0, 0.};
  int r = sums(U, thin, false);
  const double w[3][4] = {{1., 0., 0., 0.}};
  return r ? r : lincomb(U->ctx, d_out, 1, w);
}

int mag_total_rectangular_async(mag_field U, double* d_out) {
  if (!is_link(U) || !d_out) return MAG_ERR_ARG;
  mag_action thin{MAG_IWASAKI, 0., 0., 0, 0, 0.};
  int r = sums(U, thin, true);
  const double w[3][4] = {{0., 1., 0., 0.}};
  return r ? r : lincomb(U->ctx, d_out, 1, w);
}

int mag_hamiltonian_async(mag_field U, mag_field P, const mag_action* act, double* d_out) {
  if (!is_link(U) || !is_alg(P) || !act || !d_out || !same_geom(U, P) || U->ctx != P->ctx) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  int r = sums(U, *act, act->type == MAG_IWASAKI);
  if (r || (r = kinetic_sum(c, P->g, P->prec, P->p, act->mag, 2))) return r;
  // S = wp Sp + wr Sr (field-hmc.h:632-640), K = sum / 2
  const double wp = action_from_sums(*act, 1., 0.), wr = action_from_sums(*act, 0., 1.);
  const double w[3][4] = {{0., 0., 0.5, 0.}, {wp, wr, 0., 0.}, {wp, wr, 0.5, 0.}};
  return lincomb(c, d_out, 3, w);
}

int mag_avg_wilson_loop_async(mag_field U, const int r3[3], int T, double* d_out) {
  if (!is_link(U) || !r3 || T <= 0 || !d_out) return MAG_ERR_ARG;
  for (int i = 0; i < 3; ++i)
    if (r3[i] < 0) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  int nperm = 1;
  int r = dispatch_link(c, U->prec, U->kind, [&](auto t) -> int {
    return Ops<typename decltype(t)::T, decltype(t)::N>::wilson_loop(c, U->g, U->p, r3, T, 3, &nperm);
  });
  const double w[3][4] = {{0., 0., 0., 1. / (nperm * (double)U->g.V)}};
  return r ? r : lincomb(c, d_out, 1, w);
}

}  // extern "C"
