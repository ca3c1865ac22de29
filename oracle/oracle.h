// ORACLE — test infrastructure only (see oracle/su3.hpp header note).
// C ABI of the FP64 CPU restatement of the reference constrained-HMC path,
// loaded by tests/ through ctypes, by the smoke check and by the benchmark's
// CPU-baseline leg. Fields are AoS doubles in the reference's qlat layout:
// [site][mu][3][3][re,im] (144 B per link), site index
// s = x0 + L0*(x1 + L1*(x2 + L2*x3)). Momenta and forces are full 3x3
// Hermitian traceless matrices, as in the reference.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Action: type 0 = Wilson, 1 = Iwasaki form (c0 = 1 - 8 c1; DBW2 is c1 = -1.4069).
typedef struct {
  int type;
  double beta;
  double c1;
  int mag;       // constraint block size (0 = unconstrained)
  int n_stout;   // isotropic stout levels feeding the action
  double rho;    // stout weight
} orc_action;

typedef struct {
  int trajectory_length;    // MD steps per trajectory
  double dt;                // step size
  int integrator;           // 0 = force gradient, 1 = leapfrog
  int num_forced_accept;    // first trajectories accepted unconditionally
  int accept_mode;          // 0 Metropolis, 1 demon window, 2 both
  double demon_emax;        // window half-width (raw ReTr-sum units)
  uint64_t seed;            // Philox key
  int reunitarize;          // 1: reunitarize after each trajectory (reference)
} orc_hmc;

typedef struct {
  double h_old, h_new, dh, die_roll;
  int accept;
  double avg_plaq;          // of the chain state after the accept step
  double s_gauge;           // gauge action of the chain state after the step
  double demon_e[2];        // demon energies (plaquette, rectangle)
  double reunit_dev;        // max deviation reported by reunitarize
} orc_traj;

// Coupled coarse/fine md:: HMC (field-multi-md.h): rho, XI0 and md_alpha.
typedef struct {
  double rho;
  double xi0;
  double md_alpha;
} orc_multi;

typedef struct {
  double h_old, h_new, dh, die_roll;
  int accept;
  double plaq_fine, plaq_coarse;   // avg_plaquette of both fields after the step
  double parts_old[3], parts_new[3];  // kinetic, gauge, xi
  double reunit_dev[2];            // fine, coarse
} orc_multi_traj;

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_generators(double* out /* [8][18] */);
void orc_expiQ(const double* Q, double* out, long n);
void orc_DexpiQ(const double* Q, const double* DQ, double* out, long n);
void orc_compute_Lambda(const double* Q, const double* SigmaP, const double* U, double* out, long n);

void orc_haar_links(const int* L, uint64_t seed, double* U);
void orc_unit_links(const int* L, double* U);
void orc_init_momentum(const int* L, uint64_t seed, uint64_t traj, double m, double* P);
double orc_uniform(uint64_t seed, uint32_t stream, uint64_t index);

void orc_path_product(const int* L, const double* U, const int* x, const int* dir, int ndir, double* out);
void orc_staple_dagger(const int* L, const double* U, long site, int mu, double* out);
void orc_rectangular_dagger(const int* L, const double* U, long site, int mu, double* out);
void orc_chair_staple_dagger(const int* L, const double* U, long site, int mu, double* out);
// chair (twist = 0) / twist (1) path table of mu: up to max paths of 5 codes; returns the count
int orc_staple_index(int mu, int twist, int* out, int max);
void orc_twist_staple_dagger(const int* L, const double* U, long site, int mu, double* out);
int orc_is_constrained(const int* x, int mu, int mag);

double orc_total_plaq(const int* L, const double* U);
double orc_total_rectangular(const int* L, const double* U);
double orc_avg_plaquette(const int* L, const double* U);
double orc_kinetic(const int* L, const double* P, int mag);
double orc_gauge_action(const int* L, const double* U, const orc_action* a);
double orc_hamiltonian(const int* L, const double* U, const double* P, const orc_action* a, double* parts);
double orc_reunitarize(const int* L, double* U);

void orc_get_force(const int* L, const double* U, double* F, const orc_action* a);
void orc_evolve_gauge_field(const int* L, double* U, const double* P, double dt, int mag);
void orc_evolve_momentum(const int* L, double* P, const double* F, double dt, double m);
void orc_stout_smear(const int* L, const double* U, double rho, int nlev, double* out);
void orc_stout_Q(const int* L, const double* U, double rho, double* Q);
double orc_force_gradient_integrator(const int* L, double* U, double* P, const orc_action* a,
                                     int steps, double dt, int do_reunit);
void orc_leap_frog_integrator(const int* L, double* U, double* P, const orc_action* a, int steps,
                              double dt);
double orc_avg_wilson_loop(const int* L, const double* U, const int* r, int T);

int orc_run_hmc(const int* L, double* U, const orc_action* a, const orc_hmc* h, int first_traj,
                int ntraj, double* demon_state /* [4]: S0p, S0r, Ep, Er */, orc_traj* out);

void orc_demon_microcanonical(const int* L, double* U, int nsweeps, double E_max, uint64_t seed,
                              double* E /* [4] */, long* acc_rej /* [2] */);

/* mag-blocked lattices: double_multigrid coarse init (field-hmc.h:1068-1078),
 * update_constrain (field-hmc.h:988-1009), check_constrained_plaquette
 * (field-matrix.h:407-443), lattice_expansion (expansion/main.C:19-62) */
void orc_coarsen(const int* Lf, const double* Uf, int mag, double* Uc);
void orc_update_constrain(const int* Lf, double* Uf, const double* Uc, int mag);
double orc_constrained_plaquette(const int* L, const double* U, int mag);
void orc_expand(const int* Lc, const double* Uc, int mag, double* Uf);
// md:: coarse/fine HMC; Lf = fine extents (even), coarse = Lf / 2,
// xi = [coarse site][mu] doubles.
double orc_xi_energy(const int* Lf, const double* UF, const double* UC, const double* xi, const orc_multi* mp);
void orc_multi_gb(const int* Lf, const double* UF, const orc_multi* mp, double* Gb);
void orc_multi_Q(const int* Lf, const double* UF, const orc_multi* mp, double* Q);
void orc_multi_fforce(const int* Lf, const double* UF, const double* UC, const double* xi, const orc_action* a,
                      const orc_multi* mp, double* F);
void orc_multi_cforce(const int* Lf, const double* UF, const double* PF, const orc_multi* mp, double* CF);
void orc_multi_DQ(const int* Lf, const double* UF, const double* PF, const orc_multi* mp, double* DQ);
void orc_multi_heatbath(const int* Lf, double* UC, const double* UF, const orc_multi* mp, uint64_t seed,
                        uint64_t traj);
void orc_multi_zeta_T(const int* Lf, double* UF, const double* PF, double* UC, double dt, int mag,
                      const orc_multi* mp);
void orc_multi_integrate(const int* Lf, double* UF, double* PF, double* UC, const double* xi, const orc_action* a,
                         const orc_multi* mp, int which, int steps, double dt, double* dev);
double orc_multi_hamiltonian(const int* Lf, const double* UF, const double* PF, const double* UC, const double* xi,
                             const orc_action* a, const orc_multi* mp, double* parts);
int orc_run_hmc_multi(const int* Lf, double* UF, double* UC, const orc_action* a, const orc_hmc* h,
                      const orc_multi* mp, int first_traj, int ntraj, orc_multi_traj* out);
// algGCtrnHeatBath::run (alg_gchb.h:290-387): order 0 = reference site
// order, 1 = 16 checkerboard classes; acc[4] = accept, reject, acceptCons,
// rejectCons; plaq / cons_plaq (may be NULL) get one value per iteration.
void orc_local_metropolis(const int* L, double* U, double beta, int mag, int num_iter, int nhits, double small,
                          uint64_t seed, int order, long* acc, double* plaq, double* cons_plaq);
void orc_metropolis_link(double* U, const double* sigma, double beta, int nhits, double small, uint64_t seed,
                         uint64_t iter, uint32_t link, long* acc);
// Checkerboard demon with a demon per link: Elink [site][mu][4] in/out
// (plaquette, rectangle, chair, twist); 128 commuting classes per direction.
void orc_demon_checkerboard(const int* L, double* U, int nsweeps, double E_max, uint64_t seed, double* Elink,
                            long* acc_rej);
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
