// Internals of the device-resident HMC chain shared by the chain C ABI
// (chain.hip) and the serving loop (server.hip).
#pragma once
#include "runtime.hpp"

struct mag_chain_s {
  mag_ctx ctx = nullptr;      // context whose stream and workspace the chain uses
  bool owns_ctx = false;      // private child context (serving)
  int dims[4] = {0, 0, 0, 0};
  mag::Geom g;
  mag_action act{};
  mag_hmc_params p{};
  int prec = MAG_F64, kind = MAG_LINK18;
  mag_field U = nullptr, Ubak = nullptr, P = nullptr;
  // FP64 fields of the reversibility check (created on first use)
  mag_field U64 = nullptr, U64b = nullptr, P64 = nullptr;
  uint64_t* d_epoch = nullptr;  // device copy of the trajectory index (captured trajectories)
  long traj = 0;
  bool demon_init = false;
  double demon[4] = {0., 0., 0., 0.};  // S0p, S0r, Ep, Er
  double emax = 0.;
  int naccept = 0;
  // Action sums of the current links (smeared plaquette and rectangle sums,
  // thin plaquette sum), left by the previous trajectory's accept step: the
  // next H_old is evaluated on exactly those links (kept on accept, restored
  // from the backup on reject), so its sums are these values. Valid only
  // within one mag_chain_run call (the links can be written between calls).
  bool s_valid = false, s_used = false;
  double s_cache[3] = {0., 0., 0.};
};

namespace mag {

// scalar slots of one trajectory in the chain context's d_scalar
enum { S_OLD_P = 0, S_OLD_R = 1, S_OLD_K = 2, S_NEW_P = 4, S_NEW_R = 5, S_NEW_K = 6, S_DEV = 7, S_THIN_OLD = 8,
       S_THIN_NEW = 10, S_COUNT = 12 };

// All kernels of one trajectory (momentum refresh, H_old, backup, MD,
// reunitarize, H_new). With use_epoch_ptr the momentum stream reads the
// trajectory index from c->d_epoch (so a recorded trajectory can replay).
int chain_traj_kernels(mag_chain c, bool use_epoch_ptr);
// Accept step of the trajectory whose scalars are s[S_COUNT]; restores the
// backup on reject (enqueued on the chain's stream) and advances c->traj.
int chain_accept(mag_chain c, const double* s, mag_traj_result* out);
int chain_init_demon(mag_chain c);
bool chain_needs_rect(mag_chain c);

// A context sharing `parent`'s device and the given stream, with its own
// workspace, partial-sum and scalar buffers.
int ctx_child_create(mag_ctx parent, hipStream_t stream, std::shared_ptr<MemPool> pool, mag_ctx* out);
// Footprint of a chain in device bytes (fields + workspace), for admission.
size_t chain_footprint(const int dims[4], const mag_action& act, const mag_hmc_params& p);

}  // namespace mag
