// Format-independent operations: final reductions, momenta, conversions and
// the per-matrix test hooks.
#include "kernels.hpp"
#include "runtime.hpp"

namespace mag {

namespace {
template <class TS>
AlgView<TS> av(void* p, const Geom& g) { return AlgView<TS>{(TS*)p, g.V}; }
}  // namespace

int final_reduce(mag_ctx c, const double* part, int n, int stride, int op, int slot) {
  {
    KTimer t(c, "reduce");
    if (op)
      hipLaunchKernelGGL(k_final_reduce<1>, dim3(1), dim3(BLOCK), 0, c->stream, part, n, stride, c->d_scalar + slot);
    else
      hipLaunchKernelGGL(k_final_reduce<0>, dim3(1), dim3(BLOCK), 0, c->stream, part, n, stride, c->d_scalar + slot);
  }
  CHECK_LAUNCH(c);
  return MAG_OK;
}

int kinetic_sum(mag_ctx c, const Geom& g, int prec, void* P, int mag, int slot) {
  const int nb = (g.V + BLOCK - 1) / BLOCK;
  int r;
  if ((r = ensure_part(c, (size_t)nb))) return r;
  {
    KTimer t(c, "kinetic");
    if (prec == MAG_F64)
      hipLaunchKernelGGL((k_kinetic_partial<double>), dim3(nb), dim3(BLOCK), 0, c->stream, av<double>(P, g), g, mag,
                         c->d_part);
    else
      hipLaunchKernelGGL((k_kinetic_partial<float>), dim3(nb), dim3(BLOCK), 0, c->stream, av<float>(P, g), g, mag,
                         c->d_part);
  }
  CHECK_LAUNCH(c);
  return final_reduce(c, c->d_part, nb, 1, 0, slot);
}

int mom_refresh(mag_ctx c, const Geom& g, int prec, void* P, uint64_t seed, uint64_t epoch, double sig) {
  KTimer t(c, "mom_refresh");
  if (prec == MAG_F64)
    hipLaunchKernelGGL((k_mom_refresh<double>), grid_links(g), dim3(BLOCK), 0, c->stream, av<double>(P, g), g.V, seed,
                       epoch, sig);
  else
    hipLaunchKernelGGL((k_mom_refresh<float>), grid_links(g), dim3(BLOCK), 0, c->stream, av<float>(P, g), g.V, seed,
                       epoch, sig);
  CHECK_LAUNCH(c);
  return MAG_OK;
}

int mom_update(mag_ctx c, const Geom& g, int prec, void* P, void* F, double eps) {
  KTimer t(c, "mom_update");
  if (prec == MAG_F64)
    hipLaunchKernelGGL((k_mom_update<double>), grid_links(g), dim3(BLOCK), 0, c->stream, av<double>(P, g),
                       av<double>(F, g), g.V, eps);
  else
    hipLaunchKernelGGL((k_mom_update<float>), grid_links(g), dim3(BLOCK), 0, c->stream, av<float>(P, g),
                       av<float>(F, g), g.V, eps);
  CHECK_LAUNCH(c);
  return MAG_OK;
}

int alg_scale(mag_ctx c, const Geom& g, int prec, void* P, double s) {
  if (prec == MAG_F64)
    hipLaunchKernelGGL((k_alg_scale<double>), grid_links(g), dim3(BLOCK), 0, c->stream, av<double>(P, g), g.V, s);
  else
    hipLaunchKernelGGL((k_alg_scale<float>), grid_links(g), dim3(BLOCK), 0, c->stream, av<float>(P, g), g.V, s);
  CHECK_LAUNCH(c);
  return MAG_OK;
}

int convert_links(mag_ctx c, const Geom& g, int prec_in, int kind_in, void* in, int prec_out, int kind_out,
                  void* out) {
  const bool in64 = prec_in == MAG_F64 && kind_in == MAG_LINK18;
  const bool in32 = prec_in == MAG_F32 && kind_in == MAG_LINK12;
  const bool out64 = prec_out == MAG_F64 && kind_out == MAG_LINK18;
  const bool out32 = prec_out == MAG_F32 && kind_out == MAG_LINK12;
  if (in64 && out32)
    hipLaunchKernelGGL((k_convert_links<double, 18, float, 12>), grid_links(g), dim3(BLOCK), 0, c->stream,
                       LinkView<double, 18>{(double*)in, g.V}, LinkView<float, 12>{(float*)out, g.V}, g.V);
  else if (in32 && out64)
    hipLaunchKernelGGL((k_convert_links<float, 12, double, 18>), grid_links(g), dim3(BLOCK), 0, c->stream,
                       LinkView<float, 12>{(float*)in, g.V}, LinkView<double, 18>{(double*)out, g.V}, g.V);
  else
    return fail(c, MAG_ERR_UNSUPPORTED, "unsupported link conversion");
  CHECK_LAUNCH(c);
  return MAG_OK;
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
