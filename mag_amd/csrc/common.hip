// Format-independent operations: final reductions, momenta, conversions and
// the per-matrix test hooks.
#include "kernels.hpp"
#include "launch.hpp"
#include "runtime.hpp"

namespace mag {

namespace {
template <class TS>
AlgView<TS> av(void* p, const Geom& g) { return AlgView<TS>{(TS*)p, g.V}; }
}  // namespace

int final_reduce(mag_ctx c, const double* part, int n, int stride, int op, int slot) {
  ReduceArgs a{part, n, stride, c->d_scalar + slot};
  if (op) return launch(c, "reduce", k_final_reduce<1>, k_final_reduce_b<1>, dim3(1), a);
  return launch(c, "reduce", k_final_reduce<0>, k_final_reduce_b<0>, dim3(1), a);
}

int kinetic_sum(mag_ctx c, const Geom& g, int prec, void* P, int mag, int slot) {
  const int nb = (g.V + BLOCK - 1) / BLOCK;
  int r;
  if ((r = ensure_part(c, (size_t)nb))) return r;
  if (prec == MAG_F64) {
    KineticArgs<double> a{av<double>(P, g), g, mag, c->d_part};
    r = launch(c, "kinetic", k_kinetic_partial<double>, k_kinetic_partial_b<double>, dim3(nb), a);
  } else {
    KineticArgs<float> a{av<float>(P, g), g, mag, c->d_part};
    r = launch(c, "kinetic", k_kinetic_partial<float>, k_kinetic_partial_b<float>, dim3(nb), a);
  }
  if (r) return r;
  return final_reduce(c, c->d_part, nb, 1, 0, slot);
}

int mom_refresh(mag_ctx c, const Geom& g, int prec, void* P, uint64_t seed, uint64_t epoch, double sig,
                const uint64_t* epoch_ptr) {
  if (prec == MAG_F64) {
    MomArgs<double> a{av<double>(P, g), g.V, seed, epoch, epoch_ptr, sig};
    return launch(c, "mom_refresh", k_mom_refresh<double>, k_mom_refresh_b<double>, grid_links(g), a);
  }
  MomArgs<float> a{av<float>(P, g), g.V, seed, epoch, epoch_ptr, sig};
  return launch(c, "mom_refresh", k_mom_refresh<float>, k_mom_refresh_b<float>, grid_links(g), a);
}

int mom_update(mag_ctx c, const Geom& g, int prec, void* P, void* F, double eps) {
  if (prec == MAG_F64) {
    MomUpdateArgs<double> a{av<double>(P, g), av<double>(F, g), g.V, eps};
    return launch(c, "mom_update", k_mom_update<double>, k_mom_update_b<double>, grid_links(g), a);
  }
  MomUpdateArgs<float> a{av<float>(P, g), av<float>(F, g), g.V, eps};
  return launch(c, "mom_update", k_mom_update<float>, k_mom_update_b<float>, grid_links(g), a);
}

int copy_dev(mag_ctx c, void* dst, const void* src, size_t bytes) { return copy_async(c, dst, src, bytes); }

int fill_dev(mag_ctx c, double* p, double v, int n) { return fill_async(c, p, v, n); }

int alg_scale(mag_ctx c, const Geom& g, int prec, void* P, double s) {
  if (prec == MAG_F64)
    hipLaunchKernelGGL((k_alg_scale<double>), grid_links(g), dim3(BLOCK), 0, c->stream, av<double>(P, g), g.V, s);
  else
    hipLaunchKernelGGL((k_alg_scale<float>), grid_links(g), dim3(BLOCK), 0, c->stream, av<float>(P, g), g.V, s);
  CHECK_LAUNCH(c);
  return MAG_OK;
}

// Link format conversion between FP64 18-real, FP64 12-real and FP32 12-real
// (through FP64 arithmetic; 12-real outputs keep rows 0 and 1).
int convert_links(mag_ctx c, const Geom& g, int prec_in, int kind_in, void* in, int prec_out, int kind_out,
                  void* out) {
  return dispatch_link(c, prec_in, kind_in, [&](auto ti) -> int {
    using TI = typename decltype(ti)::T;
    constexpr int NI = decltype(ti)::N;
    return dispatch_link(c, prec_out, kind_out, [&](auto to) -> int {
      using TO = typename decltype(to)::T;
      constexpr int NO = decltype(to)::N;
      hipLaunchKernelGGL((k_convert_links<TI, NI, TO, NO>), grid_links(g), dim3(BLOCK), 0, c->stream,
                         LinkView<TI, NI>{(TI*)in, g.V}, LinkView<TO, NO>{(TO*)out, g.V}, g.V);
      CHECK_LAUNCH(c);
      return MAG_OK;
    });
  });
}

int batch_expiQ(mag_ctx c, const double* dQ, double* dout, long n, int prec) {
  const dim3 grid((unsigned)((n + BLOCK - 1) / BLOCK));
  if (prec == MAG_F64)
    hipLaunchKernelGGL(k_expiQ_batch<double>, grid, dim3(BLOCK), 0, c->stream, dQ, dout, n);
  else
    hipLaunchKernelGGL(k_expiQ_batch<float>, grid, dim3(BLOCK), 0, c->stream, dQ, dout, n);
  CHECK_LAUNCH(c);
  return MAG_OK;
}

int batch_dexpiq(mag_ctx c, const double* dQ, const double* dDQ, double* dout, long n, int prec) {
  const dim3 grid((unsigned)((n + BLOCK - 1) / BLOCK));
  if (prec == MAG_F64)
    hipLaunchKernelGGL(k_dexpiq_batch<double>, grid, dim3(BLOCK), 0, c->stream, dQ, dDQ, dout, n);
  else
    hipLaunchKernelGGL(k_dexpiq_batch<float>, grid, dim3(BLOCK), 0, c->stream, dQ, dDQ, dout, n);
  CHECK_LAUNCH(c);
  return MAG_OK;
}

int batch_lambda(mag_ctx c, const double* dQ, const double* dS, const double* dU, double* dout, long n, int prec) {
  const dim3 grid((unsigned)((n + BLOCK - 1) / BLOCK));
  if (prec == MAG_F64)
    hipLaunchKernelGGL(k_lambda_batch<double>, grid, dim3(BLOCK), 0, c->stream, dQ, dS, dU, dout, n);
  else
    hipLaunchKernelGGL(k_lambda_batch<float>, grid, dim3(BLOCK), 0, c->stream, dQ, dS, dU, dout, n);
  CHECK_LAUNCH(c);
  return MAG_OK;
}

// Die roll of trajectory `index` (host side; same Philox stream as the oracle).
double host_uniform(uint64_t seed, uint32_t stream, uint64_t index) {
  const u32x4 w = rng_draw(seed, stream, 0, (uint32_t)index, (uint32_t)(index >> 32));
  return rng_u01(w.x, w.y);
}

}  // namespace mag
