// mag_amd.hpp — C++ host API with the reference's entry-point names
// (hummingtree/mag field-hmc.h, field-matrix.h, field-wloop.h,
// field-demon.h, NERSC I/O of field-matrix.h, the md:: coarse/fine HMC of
// field-multi-md.h), as thin inline wrappers over
// the C ABI of mag_amd.h.
// A caller of the reference replaces qlat::Field<cps::Matrix> by
// mag::Field and keeps its call sequence:
//
//   mag::Context ctx(0);
//   mag::Field U(ctx, {32, 32, 32, 64}, MAG_LINK18, MAG_F64), P(...ALG8...);
//   mag::Arg_chmc arg; arg.trajectory_length = 12; arg.dt = 1. / 12; ...
//   mag::init_momentum(P, seed, traj);
//   std::vector<double> part;
//   double H0 = mag::get_hamiltonian(U, P, arg, part);
//   mag::force_gradient_integrator(U, P, arg);
//
// Errors are thrown as mag::Error (status code + message); the reference
// aborts through assert instead.
#pragma once
#include <array>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mag_amd.h"

namespace mag {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int code, mag_ctx ctx) {
  if (code != MAG_OK) throw Error(code, std::string("mag_amd: ") + (ctx ? mag_last_error(ctx) : "error"));
}

class Context {
 public:
  explicit Context(int device = 0) { check(mag_ctx_create(device, &h_), nullptr); }
  ~Context() {
    if (h_) mag_ctx_destroy(h_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  mag_ctx get() const { return h_; }
  void synchronize() const { check(mag_ctx_synchronize(h_), h_); }

 private:
  mag_ctx h_ = nullptr;
};

class Field {
 public:
  Field(const Context& ctx, std::array<int, 4> dims, int kind = MAG_LINK18, int prec = MAG_F64)
      : ctx_(ctx.get()), dims_(dims), kind_(kind) {
    check(mag_field_create(ctx_, dims.data(), kind, prec, &h_), ctx_);
  }
  ~Field() {
    if (h_) mag_field_destroy(h_);
  }
  Field(const Field&) = delete;
  Field& operator=(const Field&) = delete;
  // gFieldExt = gField (field-hmc.h:829)
  void copy_from(const Field& o) { check(mag_field_copy(h_, o.h_), ctx_); }
  mag_field get() const { return h_; }
  mag_ctx ctx() const { return ctx_; }
  long volume() const { return (long)dims_[0] * dims_[1] * dims_[2] * dims_[3]; }
  // qlat layout [site][mu][3][3][re,im] ([site][mu] for MAG_REAL1)
  void upload(const std::vector<double>& aos) { check(mag_field_upload_aos(h_, aos.data()), ctx_); }
  std::vector<double> download() const {
    std::vector<double> a((size_t)volume() * 4 * (kind_ == MAG_REAL1 ? 1 : 18));
    check(mag_field_download_aos(h_, a.data()), ctx_);
    return a;
  }

 private:
  mag_ctx ctx_;
  std::array<int, 4> dims_;
  int kind_;
  mag_field h_ = nullptr;
};

// Gauge{type, c1} and Arg_chmc (field-hmc.h:294-307), plus smearing.
struct Gauge {
  int type = MAG_WILSON;
  double c1 = 0.;
};
struct Arg_chmc {
  int mag = 0;
  int trajectory_length = 8;
  int num_trajectory = 1;
  int num_step_between_output = 1;
  int num_forced_accept_step = 0;
  int num_step_before_output = 0;
  double beta = 6.0;
  double dt = 0.125;
  Gauge gauge;
  int n_stout = 0;
  double rho = 0.1;
  mag_action action() const { return mag_action{gauge.type, beta, gauge.c1, mag, n_stout, rho}; }
};

// get_force (field-hmc.h:339)
inline void get_force(Field& F, const Field& U, const Arg_chmc& arg) {
  mag_action a = arg.action();
  check(mag_gauge_force(F.get(), U.get(), &a), U.ctx());
}
// evolve_momentum (field-hmc.h:427)
inline void evolve_momentum(Field& P, const Field& F, double dt, const Arg_chmc&, double m = 1.) {
  check(mag_evolve_momentum(P.get(), F.get(), dt, m), P.ctx());
}
// evolve_gauge_field (field-hmc.h:442)
inline void evolve_gauge_field(Field& U, const Field& P, double dt, const Arg_chmc& arg) {
  check(mag_evolve_gauge(U.get(), P.get(), dt, arg.mag), U.ctx());
}
// force_gradient_integrator (field-hmc.h:525); returns the reunitarize deviation
inline double force_gradient_integrator(Field& U, Field& P, const Arg_chmc& arg) {
  mag_action a = arg.action();
  double dev = 0.;
  check(mag_fg_integrate(U.get(), P.get(), &a, arg.trajectory_length, arg.dt, 1, &dev), U.ctx());
  return dev;
}
// leap_frog_integrator (field-hmc.h:569)
inline void leap_frog_integrator(Field& U, Field& P, const Arg_chmc& arg) {
  mag_action a = arg.action();
  check(mag_leapfrog_integrate(U.get(), P.get(), &a, arg.trajectory_length, arg.dt), U.ctx());
}
// get_hamiltonian (field-hmc.h:589)
inline double get_hamiltonian(const Field& U, const Field& P, const Arg_chmc& arg, std::vector<double>& part) {
  mag_action a = arg.action();
  double parts[2], H = 0.;
  check(mag_hamiltonian(U.get(), P.get(), &a, parts, &H), U.ctx());
  part.assign(parts, parts + 2);
  return H;
}
// get_hamiltonian left on the device: d_out[0..2] = {kinetic, potential, H}
inline void get_hamiltonian_async(const Field& U, const Field& P, const Arg_chmc& arg, double* d_out) {
  mag_action a = arg.action();
  check(mag_hamiltonian_async(U.get(), P.get(), &a, d_out), U.ctx());
}
// init_momentum (field-hmc.h:646), Philox stream keyed by (seed, trajectory)
inline void init_momentum(Field& P, uint64_t seed, uint64_t traj, double m = 1.) {
  check(mag_init_momentum(P.get(), seed, traj, m), P.ctx());
}
// total_plaq / total_rectangular / avg_plaquette (field-matrix.h:299, 354, 387)
inline double total_plaq(const Field& U) {
  double v = 0.;
  check(mag_total_plaq(U.get(), &v), U.ctx());
  return v;
}
inline double total_rectangular(const Field& U) {
  double v = 0.;
  check(mag_total_rectangular(U.get(), &v), U.ctx());
  return v;
}
inline double avg_plaquette(const Field& U) {
  double v = 0.;
  check(mag_avg_plaquette(U.get(), &v), U.ctx());
  return v;
}
// Asynchronous variants: results stay in device memory (d_out), written in
// the context's stream order.
inline void total_plaq_async(const Field& U, double* d_out) { check(mag_total_plaq_async(U.get(), d_out), U.ctx()); }
inline void total_rectangular_async(const Field& U, double* d_out) {
  check(mag_total_rectangular_async(U.get(), d_out), U.ctx());
}
inline void avg_wilson_loop_async(const Field& U, const std::array<int, 3>& r, int T, double* d_out) {
  check(mag_avg_wilson_loop_async(U.get(), r.data(), T, d_out), U.ctx());
}
// reunitarize (field-matrix.h:83)
inline double reunitarize(Field& U) {
  double v = 0.;
  check(mag_reunitarize(U.get(), &v), U.ctx());
  return v;
}
// avg_wilson_loop (field-wloop.h:58)
inline double avg_wilson_loop(const Field& U, const std::array<int, 3>& r, int T) {
  double v = 0.;
  check(mag_avg_wilson_loop(U.get(), r.data(), T, &v), U.ctx());
  return v;
}
// demon_microcanonical (field-demon.h:25); E holds the 4 demon energies
inline std::pair<long, long> demon_microcanonical(Field& U, int nsweeps, double E_max, uint64_t seed,
                                                  std::array<double, 4>& E) {
  long ar[2] = {0, 0};
  check(mag_demon_microcanonical(U.get(), nsweeps, E_max, seed, E.data(), ar), U.ctx());
  return {ar[0], ar[1]};
}

// ------------------------------------------------- mag-blocked lattices
// coarse links as products of mag fine links (double_multigrid's coarse
// initialisation, field-hmc.h:1068-1078)
inline void coarsen_links(const Field& fine, int mag, Field& coarse) {
  check(mag_coarsen_links(fine.get(), mag, coarse.get()), fine.ctx());
}
// update_constrain (field-hmc.h:988)
inline void update_constrain(Field& gField, const Field& gField_coarse, const Arg_chmc& arg) {
  check(mag_update_constrain(gField.get(), gField_coarse.get(), arg.mag), gField.ctx());
}
// check_constrained_plaquette (field-matrix.h:407)
inline double check_constrained_plaquette(const Field& gField, int mag) {
  double v = 0.;
  check(mag_constrained_plaquette(gField.get(), mag, &v), gField.ctx());
  return v;
}

// Arg_export (field-matrix.h:445-450)
struct Arg_export {
  double beta = 0.;
  int sequence_num = 0;
  std::string ensemble_label;
};
// export_config_nersc (field-matrix.h:452)
inline void export_config_nersc(const Field& U, const std::string& path, const Arg_export& arg,
                                bool doesSkipThird = false) {
  check(mag_export_config_nersc(U.get(), path.c_str(), arg.beta, arg.sequence_num, arg.ensemble_label.c_str(),
                                doesSkipThird ? 1 : 0),
        U.ctx());
}
// import_config_nersc (field-matrix.h:592)
inline mag_nersc_header import_config_nersc(Field& U, const std::string& path) {
  mag_nersc_header h{};
  check(mag_import_config_nersc(U.get(), path.c_str(), &h), U.ctx());
  return h;
}

// algGCtrnHeatBath (alg_gchb.h:107-387): gchbArg plus the Wilson beta (GJP.Beta() in CPS)
struct gchbArg {
  int mag = 0;
  int numIter = 1;
  int nHits = 1;
  double small = 0.5;
};
struct LocalResult {
  long accept = 0, reject = 0, acceptCons = 0, rejectCons = 0;
  std::vector<double> avgPlaq, avgConsPlaq;  // per iteration
};
inline LocalResult local_metropolis(Field& U, double beta, const gchbArg& arg, uint64_t seed,
                                    int order = MAG_ORDER_CHECKERBOARD) {
  LocalResult r;
  long a[4] = {0, 0, 0, 0};
  r.avgPlaq.assign((size_t)arg.numIter, 0.);
  r.avgConsPlaq.assign((size_t)arg.numIter, 0.);
  check(mag_local_metropolis(U.get(), beta, arg.mag, arg.numIter, arg.nHits, arg.small, seed, order, a,
                             r.avgPlaq.data(), r.avgConsPlaq.data()),
        U.ctx());
  r.accept = a[0];
  r.reject = a[1];
  r.acceptCons = a[2];
  r.rejectCons = a[3];
  return r;
}
// demon_microcanonical with a demon per link (MAG_REAL1 fields of U's shape)
inline std::pair<long, long> demon_checkerboard(Field& U, Field& Ep, Field& Er, Field& Ec, Field& Et, int nsweeps,
                                                double E_max, uint64_t seed) {
  long ar[2] = {0, 0};
  check(mag_demon_checkerboard(U.get(), Ep.get(), Er.get(), Ec.get(), Et.get(), nsweeps, E_max, seed, ar), U.ctx());
  return {ar[0], ar[1]};
}

// run_chmc (field-hmc.h:770) as a device-resident chain.
class Chain {
 public:
  Chain(const Context& ctx, std::array<int, 4> dims, const Arg_chmc& arg, const mag_hmc_params& p)
      : ctx_(ctx.get()) {
    mag_action a = arg.action();
    check(mag_chain_create(ctx_, dims.data(), &a, &p, &h_), ctx_);
  }
  ~Chain() {
    if (h_) mag_chain_destroy(h_);
  }
  Chain(const Chain&) = delete;
  Chain& operator=(const Chain&) = delete;
  void set_links(const std::vector<double>& aos) { check(mag_chain_set_links_aos(h_, aos.data()), ctx_); }
  void haar(uint64_t seed) { check(mag_chain_haar_links(h_, seed), ctx_); }
  std::vector<mag_traj_result> run(int ntraj) {
    std::vector<mag_traj_result> r((size_t)ntraj);
    check(mag_chain_run(h_, ntraj, r.data()), ctx_);
    return r;
  }
  mag_chain get() const { return h_; }

 private:
  mag_ctx ctx_;
  mag_chain h_ = nullptr;
};

// The coupled coarse/fine HMC of field-multi-md.h (namespace md:: there too).
// Fine fields have even extents, coarse fields half of them; CxField is a
// MAG_REAL1 field of the coarse shape. rho and XI0 are the includer's
// constants in the reference; here they travel in mag_multi_params.
namespace md {

constexpr double md_alpha = 0.5;  // field-multi-md.h:28

inline double get_xi_energy(const Field& FgField, const Field& CgField, const Field& CxField,
                            const mag_multi_params& mp) {
  double v = 0.;
  check(mag_xi_energy(FgField.get(), CgField.get(), CxField.get(), &mp, &v), FgField.ctx());
  return v;
}
inline double get_hamiltonian_multi(const Field& FgField, const Field& FmField, const Arg_chmc& Farg,
                                    const Field& CgField, const Field& CxField, const mag_multi_params& mp,
                                    std::vector<double>& part) {
  mag_action a = Farg.action();
  double h = 0.;
  part.assign(3, 0.);
  check(mag_multi_hamiltonian(FgField.get(), FmField.get(), CgField.get(), CxField.get(), &a, &mp, part.data(), &h),
        FgField.ctx());
  return h;
}
inline void get_Fforce(Field& FfField, const Field& FgField, const Field& CgField, const Arg_chmc& Farg,
                       const Field& CxField, const mag_multi_params& mp) {
  mag_action a = Farg.action();
  check(mag_multi_fforce(FfField.get(), FgField.get(), CgField.get(), CxField.get(), &a, &mp), FgField.ctx());
}
inline void get_Cforce(Field& CfField, const Field& FgField, const Field& FmField, const mag_multi_params& mp) {
  check(mag_multi_cforce(CfField.get(), FgField.get(), FmField.get(), &mp), FgField.ctx());
}
inline void get_Gb(Field& CGbField, const Field& FgField, const mag_multi_params& mp) {
  check(mag_multi_gb(CGbField.get(), FgField.get(), &mp), FgField.ctx());
}
inline void heatbath(Field& CgField, const Field& FgField, const mag_multi_params& mp, uint64_t seed,
                     uint64_t traj) {
  check(mag_multi_heatbath(CgField.get(), FgField.get(), &mp, seed, traj), FgField.ctx());
}
inline void force_gradient_integrator_zeta_T(Field& FgField, const Field& FmField, Field& CgField, double dt,
                                             const Arg_chmc& Farg, const mag_multi_params& mp) {
  check(mag_multi_zeta_T(FgField.get(), FmField.get(), CgField.get(), dt, Farg.mag, &mp), FgField.ctx());
}
// returns the reunitarize deviations {fine, coarse}
inline std::array<double, 2> force_gradient_integrator_zeta_H(Field& FgField, Field& FmField, Field& CgField,
                                                              const Field& CxField, const Arg_chmc& Farg,
                                                              const mag_multi_params& mp) {
  mag_action a = Farg.action();
  std::array<double, 2> dev{};
  check(mag_multi_integrate(FgField.get(), FmField.get(), CgField.get(), CxField.get(), &a, &mp, MAG_MULTI_ZETA_H,
                            Farg.trajectory_length, Farg.dt, dev.data()),
        FgField.ctx());
  return dev;
}
inline std::array<double, 2> force_gradient_integrator(Field& FgField, Field& FmField, const Arg_chmc& Farg,
                                                       int steps, Field& CgField, const Field& CxField,
                                                       const mag_multi_params& mp) {
  mag_action a = Farg.action();
  std::array<double, 2> dev{};
  check(mag_multi_integrate(FgField.get(), FmField.get(), CgField.get(), CxField.get(), &a, &mp, MAG_MULTI_FG, steps,
                            Farg.dt, dev.data()),
        FgField.ctx());
  return dev;
}
inline void nested_integrator(Field& FgField, Field& FmField, const Arg_chmc& Farg, Field& CgField,
                              const Field& CxField, const mag_multi_params& mp) {
  mag_action a = Farg.action();
  check(mag_multi_integrate(FgField.get(), FmField.get(), CgField.get(), CxField.get(), &a, &mp, MAG_MULTI_NESTED,
                            Farg.trajectory_length, Farg.dt, nullptr),
        FgField.ctx());
}

// run_hmc_multi (field-multi-md.h:785) as a device-resident chain.
class MultiChain {
 public:
  MultiChain(const Context& ctx, std::array<int, 4> fine_dims, const Arg_chmc& Farg, const mag_hmc_params& p,
             const mag_multi_params& mp)
      : ctx_(ctx.get()) {
    mag_action a = Farg.action();
    check(mag_multi_create(ctx_, fine_dims.data(), &a, &p, &mp, &h_), ctx_);
  }
  ~MultiChain() {
    if (h_) mag_multi_destroy(h_);
  }
  MultiChain(const MultiChain&) = delete;
  MultiChain& operator=(const MultiChain&) = delete;
  // borrowed device fields (fine, coarse)
  std::pair<mag_field, mag_field> fields() const {
    mag_field f = nullptr, c = nullptr;
    check(mag_multi_fields(h_, &f, &c), ctx_);
    return {f, c};
  }
  std::vector<mag_multi_result> run(int ntraj) {
    std::vector<mag_multi_result> r((size_t)ntraj);
    check(mag_multi_run(h_, ntraj, r.data()), ctx_);
    return r;
  }

 private:
  mag_ctx ctx_;
  mag_multi h_ = nullptr;
};

}  // namespace md

}  // namespace mag
