// Mag-blocked lattices: the coarse/fine tools around the constrained HMC.
// A coarse lattice of extents L/mag has one link per mag consecutive fine
// links along a direction; the constrained fine HMC (link classes 10, 1, 100
// of is_constrained, field-hmc.h:309-337) keeps those products fixed.
//   coarsen         coarse link = ordered product of the mag fine links from
//                   mag*xC (the coarse initialisation of double_multigrid,
//                   field-hmc.h:1068-1078)
//   update_constrain  first fine link := U_c(xC) [U(x+mu) ... U(x+(mag-1)mu)]^+
//                   (field-hmc.h:988-1009), so the product equals U_c again
//   constrained plaquette  Re Tr / 3 of the mag x mag loops at the coarse
//                   sites, planes nu < mu (check_constrained_plaquette,
//                   field-matrix.h:407-443)
//   lattice expansion  unit links, the coarse link at fine site mag*x
//                   (expansion/main.C:19-62, without the file I/O)
// One thread per coarse link (or per fine link for the expansion); products
// in the field's precision, the loop traces accumulated in FP64.
#include "kernels.hpp"
#include "runtime.hpp"

using namespace mag;

namespace {

MAG_D int site_index(const Geom& g, const int* x) {
  int y[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    y[i] = x[i] % g.L[i];
    if (y[i] < 0) y[i] += g.L[i];
  }
  return y[0] + g.S[1] * y[1] + g.S[2] * y[2] + g.S[3] * y[3];
}

MAG_D void shift(int* y, int mu, int d) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i == mu) y[i] += d;
}

template <class TS, int NC>
__global__ __launch_bounds__(BLOCK) void k_coarsen(LinkView<TS, NC> uf, Geom gf, LinkView<TS, NC> uc, Geom gc,
                                                   int mag) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= gc.V) return;
  const int mu = blockIdx.y;
  const Site pc = site_of(gc, s);
  int y[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) y[i] = mag * pc.x[i];
  M3<TS> m = unit3<TS>();
  for (int k = 0; k < mag; ++k) {
    m = mul(m, uf.template load<TS>(mu, site_index(gf, y)));
    shift(y, mu, 1);
  }
  uc.store(mu, s, m);
}

template <class TS, int NC>
__global__ __launch_bounds__(BLOCK) void k_update_constrain(LinkView<TS, NC> uf, Geom gf, LinkView<TS, NC> uc,
                                                            Geom gc, int mag) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= gc.V) return;
  const int mu = blockIdx.y;
  const Site pc = site_of(gc, s);
  int x[4], y[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) y[i] = x[i] = mag * pc.x[i];
  shift(y, mu, 1);
  M3<TS> m = unit3<TS>();
  for (int k = 1; k < mag; ++k) {
    m = mul(m, uf.template load<TS>(mu, site_index(gf, y)));
    shift(y, mu, 1);
  }
  uf.store(mu, site_index(gf, x), mul_nd(uc.template load<TS>(mu, s), m));
}

template <class TS, int NC>
__global__ __launch_bounds__(BLOCK) void k_constrained_plaq(LinkView<TS, NC> u, Geom g, Geom gc, int mag,
                                                            double* part) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  double acc = 0.;
  if (s < gc.V) {
    const Site pc = site_of(gc, s);
    int x[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = mag * pc.x[i];
    for (int mu = 1; mu < 4; ++mu)
      for (int nu = 0; nu < mu; ++nu) {
        // mag steps along mu, mag along nu, mag back along mu, mag back along nu
        int y[4] = {x[0], x[1], x[2], x[3]};
        M3<double> m = unit3<double>();
        for (int k = 0; k < mag; ++k) {
          m = mul(m, u.template load<double>(mu, site_index(g, y)));
          shift(y, mu, 1);
        }
        for (int k = 0; k < mag; ++k) {
          m = mul(m, u.template load<double>(nu, site_index(g, y)));
          shift(y, nu, 1);
        }
        for (int k = 0; k < mag; ++k) {
          shift(y, mu, -1);
          m = mul_nd(m, u.template load<double>(mu, site_index(g, y)));
        }
        for (int k = 0; k < mag; ++k) {
          shift(y, nu, -1);
          m = mul_nd(m, u.template load<double>(nu, site_index(g, y)));
        }
        acc += m(0, 0).r + m(1, 1).r + m(2, 2).r;
      }
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <class TS, int NC>
__global__ __launch_bounds__(BLOCK) void k_expand(LinkView<TS, NC> uc, Geom gc, LinkView<TS, NC> uf, Geom gf,
                                                  int mag) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= gf.V) return;
  const int mu = blockIdx.y;
  const Site p = site_of(gf, s);
  bool origin = true;
  int xc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    origin = origin && (p.x[i] % mag == 0);
    xc[i] = p.x[i] / mag;
  }
  if (origin) uf.store(mu, s, uc.template load<TS>(mu, site_index(gc, xc)));
  else uf.store(mu, s, unit3<TS>());
}

bool linkf(mag_field f) { return f && (f->kind == MAG_LINK18 || f->kind == MAG_LINK12); }

// fine = mag x coarse in every direction, same format and context
int check_pair(mag_field fine, mag_field coarse, int mag) {
  if (!linkf(fine) || !linkf(coarse) || mag < 1) return MAG_ERR_ARG;
  if (fine->ctx != coarse->ctx || fine->prec != coarse->prec || fine->kind != coarse->kind)
    return fail(fine->ctx, MAG_ERR_ARG, "fine and coarse fields differ in context or format");
  for (int i = 0; i < 4; ++i)
    if (fine->dims[i] != mag * coarse->dims[i])
      return fail(fine->ctx, MAG_ERR_GEOMETRY, "fine extents must be mag times the coarse extents");
  return MAG_OK;
}

dim3 grid_of(const Geom& g) { return dim3((g.V + BLOCK - 1) / BLOCK, 4, 1); }

}  // namespace

namespace mag {
// Sum of Re Tr of the mag x mag loops at the coarse sites into scalar slot
// `slot` (no host read); *norm turns it into the constrained plaquette.
int constrained_plaq_sum(mag_ctx c, const Geom& g, int prec, int kind, void* U, int mag, int slot, double* norm) {
  int dc[4];
  for (int i = 0; i < 4; ++i) dc[i] = g.L[i] / mag;
  const Geom gc = make_geom(dc);
  const int nb = (gc.V + BLOCK - 1) / BLOCK;
  int r;
  if ((r = ensure_part(c, (size_t)nb))) return r;
  r = dispatch_link(c, prec, kind, [&](auto t) -> int {
    using TS = typename decltype(t)::T;
    constexpr int N = decltype(t)::N;
    hipLaunchKernelGGL((k_constrained_plaq<TS, N>), dim3(nb), dim3(BLOCK), 0, c->stream,
                       LinkView<TS, N>{(TS*)U, g.V}, g, gc, mag, c->d_part);
    CHECK_LAUNCH(c);
    return MAG_OK;
  });
  if (r) return r;
  *norm = 1. / (3. * 6. * (double)gc.V);
  return final_reduce(c, c->d_part, nb, 1, 0, slot);
}
}  // namespace mag

extern "C" {

int mag_coarsen_links(mag_field fine, int mag, mag_field coarse) {
  int r = check_pair(fine, coarse, mag);
  if (r) return r;
  mag_ctx c = fine->ctx;
  return dispatch_link(c, fine->prec, fine->kind, [&](auto t) -> int {
    using TS = typename decltype(t)::T;
    constexpr int N = decltype(t)::N;
    hipLaunchKernelGGL((k_coarsen<TS, N>), grid_of(coarse->g), dim3(BLOCK), 0, c->stream,
                       LinkView<TS, N>{(TS*)fine->p, fine->g.V}, fine->g, LinkView<TS, N>{(TS*)coarse->p, coarse->g.V},
                       coarse->g, mag);
    CHECK_LAUNCH(c);
    return MAG_OK;
  });
}

int mag_update_constrain(mag_field fine, mag_field coarse, int mag) {
  int r = check_pair(fine, coarse, mag);
  if (r) return r;
  mag_ctx c = fine->ctx;
  return dispatch_link(c, fine->prec, fine->kind, [&](auto t) -> int {
    using TS = typename decltype(t)::T;
    constexpr int N = decltype(t)::N;
    hipLaunchKernelGGL((k_update_constrain<TS, N>), grid_of(coarse->g), dim3(BLOCK), 0, c->stream,
                       LinkView<TS, N>{(TS*)fine->p, fine->g.V}, fine->g, LinkView<TS, N>{(TS*)coarse->p, coarse->g.V},
                       coarse->g, mag);
    CHECK_LAUNCH(c);
    return MAG_OK;
  });
}

int mag_constrained_plaquette(mag_field U, int mag, double* out) {
  if (!linkf(U) || mag < 1) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  for (int i = 0; i < 4; ++i)
    if (U->dims[i] % mag) return fail(c, MAG_ERR_GEOMETRY, "extents must be multiples of mag");
  double norm = 0.;
  int r = mag::constrained_plaq_sum(c, U->g, U->prec, U->kind, U->p, mag, 0, &norm);
  if (r || (r = fetch_scalars(c, 1))) return r;
  if (out) *out = c->h_scalar[0] * norm;
  return MAG_OK;
}

int mag_lattice_expansion(mag_field coarse, int mag, mag_field fine) {
  int r = check_pair(fine, coarse, mag);
  if (r) return r;
  mag_ctx c = fine->ctx;
  return dispatch_link(c, fine->prec, fine->kind, [&](auto t) -> int {
    using TS = typename decltype(t)::T;
    constexpr int N = decltype(t)::N;
    hipLaunchKernelGGL((k_expand<TS, N>), grid_of(fine->g), dim3(BLOCK), 0, c->stream,
                       LinkView<TS, N>{(TS*)coarse->p, coarse->g.V}, coarse->g, LinkView<TS, N>{(TS*)fine->p, fine->g.V},
                       fine->g, mag);
    CHECK_LAUNCH(c);
    return MAG_OK;
  });
}

}  // extern "C"
