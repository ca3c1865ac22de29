// Allocation-free Hamiltonian (mag_hamiltonian_noalloc): the reductions of
// mag_hamiltonian_async for a thin action, with their partial sums in a
// fixed device array of this library instead of the context's workspace, so
// a call allocates nothing (callers under graph capture or in loops that may
// not call hipMalloc). The array exists once per GPU; its device address is
// looked up once per device when a context is created on it.
#include <mutex>
#include <string>

#include "kernels.hpp"
#include "runtime.hpp"

using namespace mag;

namespace {

constexpr long NOALLOC_MAX_SITES = MAG_NOALLOC_MAX_SITES;
constexpr int NOALLOC_MAX_BLOCKS = (int)((NOALLOC_MAX_SITES + BLOCK - 1) / BLOCK);
// Layout (doubles): [0, 2B) plaquette/rectangle partial pairs of the action
// kernel, [2B, 3B) kinetic partials, [3B, 3B+4) the three sums and a slot
// that stays zero (the fourth operand of the linear combination).
constexpr size_t PART_ACTION = 0;
constexpr size_t PART_KINETIC = 2 * (size_t)NOALLOC_MAX_BLOCKS;
constexpr size_t SUMS = 3 * (size_t)NOALLOC_MAX_BLOCKS;
constexpr size_t NOALLOC_DOUBLES = SUMS + 4;

// Device scratch of the allocation-free entry point (1.5 MiB per GPU, zero at
// load). Host code must not pass `scratch` itself to a kernel: its device
// address comes from hipGetSymbolAddress (scratch_ptr below).
__device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_scratch[MAX_DEVICES] = {};

}  // namespace

namespace mag {

// Called by mag_ctx_create with `device` current: resolves this device's
// copy of the scratch array once.
int noalloc_setup(int device, std::string* err) {
  if (device < 0 || device >= MAX_DEVICES) {
    *err = "device index beyond the allocation-free scratch table";
    return MAG_ERR_DEVICE;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return MAG_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return MAG_ERR_DEVICE;
  }
  g_scratch[device] = (double*)p;
  return MAG_OK;
}

}  // namespace mag

namespace {

double* scratch_ptr(int device) {
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_scratch[device];
}

template <class TS, int NC>
int launch_action(mag_ctx c, const Geom& g, void* U, int rect, double* part) {
  const int nb = (g.V + BLOCK - 1) / BLOCK;
  ActionArgs<TS, NC> aa{LinkView<TS, NC>{(TS*)U, g.V}, g, part, rect};
  hipLaunchKernelGGL((k_action_partial<TS, NC>), dim3(nb), dim3(BLOCK), 0, c->stream, aa);
  HIPCHK(c, hipGetLastError());
  return MAG_OK;
}

template <class TS>
int launch_kinetic(mag_ctx c, const Geom& g, void* P, int mag, double* part) {
  const int nb = (g.V + BLOCK - 1) / BLOCK;
  KineticArgs<TS> ka{AlgView<TS>{(TS*)P, g.V}, g, mag, part};
  hipLaunchKernelGGL((k_kinetic_partial<TS>), dim3(nb), dim3(BLOCK), 0, c->stream, ka);
  HIPCHK(c, hipGetLastError());
  return MAG_OK;
}

int launch_sum(mag_ctx c, const double* part, int n, int stride, double* out) {
  ReduceArgs ra{part, n, stride, out};
  hipLaunchKernelGGL((k_final_reduce<0>), dim3(1), dim3(BLOCK), 0, c->stream, ra);
  HIPCHK(c, hipGetLastError());
  return MAG_OK;
}

}  // namespace

extern "C" {

int mag_hamiltonian_noalloc(mag_field U, mag_field P, const mag_action* act, double* d_out) {
  if (!U || !P || !act || !d_out || U->ctx != P->ctx) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  if ((U->kind != MAG_LINK18 && U->kind != MAG_LINK12) || P->kind != MAG_ALG8)
    return fail(c, MAG_ERR_ARG, "hamiltonian_noalloc: U must be a link field and P an algebra field");
  for (int i = 0; i < 4; ++i)
    if (U->dims[i] != P->dims[i]) return fail(c, MAG_ERR_ARG, "hamiltonian_noalloc: U and P differ in shape");
  if (act->n_stout != 0)
    return fail(c, MAG_ERR_UNSUPPORTED, "hamiltonian_noalloc: thin actions only (n_stout = 0); a smeared action "
                                        "needs smeared-field workspace");
  const Geom& g = U->g;
  if ((long)g.V > NOALLOC_MAX_SITES)
    return fail(c, MAG_ERR_ARG, "hamiltonian_noalloc: " + std::to_string(g.V) + " sites exceed the limit of " +
                                    std::to_string(NOALLOC_MAX_SITES));
  double* s = scratch_ptr(c->device);
  if (!s) return fail(c, MAG_ERR_STATE, "hamiltonian_noalloc: scratch of this device not set up");
  const int nb = (g.V + BLOCK - 1) / BLOCK;
  const int rect = act->type == MAG_IWASAKI ? 1 : 0;
  // the same kernels, grids and summation order as mag_hamiltonian_async
  int r = (U->prec == MAG_F64 && U->kind == MAG_LINK18) ? launch_action<double, 18>(c, g, U->p, rect, s + PART_ACTION)
          : (U->prec == MAG_F32 && U->kind == MAG_LINK12) ? launch_action<float, 12>(c, g, U->p, rect, s + PART_ACTION)
                                                           : fail(c, MAG_ERR_UNSUPPORTED, "unsupported link format");
  if (r || (r = launch_sum(c, s + PART_ACTION, nb, 2, s + SUMS)) ||
      (r = launch_sum(c, s + PART_ACTION + 1, nb, 2, s + SUMS + 1)))
    return r;
  r = P->prec == MAG_F64 ? launch_kinetic<double>(c, g, P->p, act->mag, s + PART_KINETIC)
                         : launch_kinetic<float>(c, g, P->p, act->mag, s + PART_KINETIC);
  if (r || (r = launch_sum(c, s + PART_KINETIC, nb, 1, s + SUMS + 2))) return r;
  // S = wp Sp + wr Sr (field-hmc.h:632-640), K = sum / 2; slot SUMS+3 is never written
  const double wp = action_from_sums(*act, 1., 0.), wr = action_from_sums(*act, 0., 1.);
  LinCombArgs a{s + SUMS, d_out, 3, {{0., 0., 0.5, 0.}, {wp, wr, 0., 0.}, {wp, wr, 0.5, 0.}}};
  hipLaunchKernelGGL(k_lincomb<0>, dim3(1), dim3(64), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  return MAG_OK;
}

}  // extern "C"
