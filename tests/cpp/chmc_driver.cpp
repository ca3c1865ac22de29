// A reference-style driver through the C++ host API (include/mag_amd.hpp):
// the call sequence of run_chmc (field-hmc.h:806-889) on a small Iwasaki
// + stout lattice, FP64, printing the per-trajectory summary row of the
// reference (field-hmc.h:862-864). Exit status 0 when every step succeeded
// and the Hamiltonian stayed finite.
#include <cmath>
#include <cstdio>
#include <vector>

#include "mag_amd.hpp"

int main(int argc, char** argv) {
  const int ntraj = argc > 1 ? std::atoi(argv[1]) : 3;
  mag::Context ctx(0);
  const std::array<int, 4> dims = {4, 4, 4, 8};
  mag::Arg_chmc arg;
  arg.trajectory_length = 4;
  arg.dt = 0.25;
  arg.beta = 2.6;
  arg.gauge.type = MAG_IWASAKI;
  arg.gauge.c1 = -0.331;
  arg.n_stout = 2;
  arg.rho = 0.1;
  arg.num_forced_accept_step = 1;
  mag::Field U(ctx, dims, MAG_LINK18, MAG_F64), Ubak(ctx, dims, MAG_LINK18, MAG_F64), P(ctx, dims, MAG_ALG8, MAG_F64);
  mag::check(mag_haar_links(U.get(), 1234), ctx.get());
  int naccept = 0;
  for (int i = 0; i < ntraj; ++i) {
    mag::init_momentum(P, 1234, i);
    std::vector<double> part_old, part_new;
    const double h_old = mag::get_hamiltonian(U, P, arg, part_old);
    Ubak.copy_from(U);
    const double dev = mag::force_gradient_integrator(U, P, arg);
    const double h_new = mag::get_hamiltonian(U, P, arg, part_new);
    const double prob = std::exp(h_old - h_new);
    const bool acc = i < arg.num_forced_accept_step || 0.5 < prob;
    if (acc) ++naccept;
    else U.copy_from(Ubak);
    std::printf("%i\t%.6e\t%.6e\t%.12e\t%i\t%.12e\treunit=%.2e\n", i + 1, std::fabs(h_new - h_old), prob,
                mag::avg_plaquette(U), acc ? 1 : 0, acc ? part_new[1] : part_old[1], dev);
    if (!std::isfinite(h_new)) return 1;
  }
  std::array<double, 4> E = {0, 0, 0, 0};
  auto ar = mag::demon_microcanonical(U, 1, 50.0, 7, E);
  std::printf("demon: accept %ld reject %ld E = %.6f %.6f %.6f %.6f\n", ar.first, ar.second, E[0], E[1], E[2], E[3]);
  std::printf("Accept Rate = %.4f\n", (double)naccept / ntraj);
  return 0;
}
