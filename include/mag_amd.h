/*
 * mag_amd — MI355X-native constrained hybrid Monte Carlo for SU(3) lattice
 * gauge fields: the C ABI of libmag_amd.so.
 *
 * The reference (hummingtree/mag) exposes its hot path as namespace-level
 * inline C++ over qlat::Field<cps::Matrix>; it has no FFI. This ABI is the
 * boundary a drop-in needs: every entry point below names the reference
 * function it replaces (file:line). Conventions:
 *   - handle based; plain pointers and sizes; no C++ or torch types;
 *   - every call returns 0 (MAG_OK) or a negative status; mag_last_error()
 *     gives the message of the last failure on a context;
 *   - fields are allocated by the caller through mag_field_create and live on
 *     the context's device; kernels never allocate on the hot path (the
 *     context owns all workspace, created on first use for a lattice shape);
 *   - host interop uses the reference's qlat layout: AoS doubles
 *     [site][mu][3][3][re,im], site = x0 + L0 (x1 + L1 (x2 + L2 x3));
 *     algebra fields (momenta, forces) are passed as Hermitian traceless 3x3
 *     matrices, as the reference stores them;
 *   - calls returning an FP64 scalar synchronise the context's stream.
 * One context per host thread at a time; contexts share nothing.
 */
#ifndef MAG_AMD_H
#define MAG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mag_ctx_s* mag_ctx;
typedef struct mag_field_s* mag_field;
typedef struct mag_chain_s* mag_chain;
typedef struct mag_server_s* mag_server;
typedef struct mag_multi_s* mag_multi;

/* status codes */
enum {
  MAG_OK = 0,
  MAG_ERR_ARG = -1,         /* bad argument or mismatched fields */
  MAG_ERR_GEOMETRY = -2,    /* lattice extents not supported */
  MAG_ERR_OOM = -3,         /* device allocation failed */
  MAG_ERR_DEVICE = -4,      /* HIP runtime error */
  MAG_ERR_NAN = -5,         /* non-finite Hamiltonian: trajectory rejected */
  MAG_ERR_UNSUPPORTED = -6, /* combination not implemented */
  MAG_ERR_BUDGET = -7,      /* chain does not fit the HBM budget */
  MAG_ERR_STATE = -8,       /* call not valid in the object's current state */
  MAG_ERR_IO = -9,          /* file missing, unreadable or not in the expected format */
  MAG_ERR_CHECKSUM = -10    /* stored checksum or plaquette does not match the data */
};

/* field kinds and precisions; MAG_REAL1 holds one real per link (the xi
 * field of the coarse/fine HMC, qlat Field<double> of multiplicity 4, host
 * layout [site][mu]) and is FP64 only */
enum { MAG_LINK18 = 0, MAG_LINK12 = 1, MAG_ALG8 = 2, MAG_REAL1 = 3 };
enum { MAG_F32 = 0, MAG_F64 = 1 };
/* gauge actions: Wilson, or the Iwasaki form with c0 = 1 - 8 c1 (Iwasaki
 * c1 = -0.331, DBW2 c1 = -1.4069); Gauge{type, c1}, field-hmc.h:344, 384 */
enum { MAG_WILSON = 0, MAG_IWASAKI = 1 };

typedef struct {
  int type;      /* MAG_WILSON or MAG_IWASAKI */
  double beta;
  double c1;     /* rectangle coefficient (Iwasaki form) */
  int mag;       /* constraint block size, 0 = unconstrained (Arg_chmc::mag) */
  int n_stout;   /* isotropic stout levels feeding the action (0 = thin) */
  double rho;    /* stout weight */
} mag_action;

/* Arg_chmc (field-hmc.h:294-307) plus the build's run controls. */
typedef struct {
  int trajectory_length;   /* MD steps per trajectory */
  double dt;               /* step size */
  int integrator;          /* 0 force gradient (field-hmc.h:525), 1 leapfrog (:569) */
  int num_forced_accept;   /* num_forced_accept_step */
  int accept_mode;         /* 0 Metropolis, 1 demon window, 2 both */
  double demon_emax;       /* > 0: window half-width; < 0: fraction of |S0| */
  uint64_t seed;           /* Philox key of the chain */
  int reunitarize;         /* 1: reunitarize after every trajectory */
  int precision;           /* MAG_F64: all FP64; MAG_F32: FP32 MD, FP64 H */
  int recon12;             /* FP64 links stored as 12 reals (rows 0, 1; row 2 rebuilt on load);
                              FP32 links are always 12-real */
} mag_hmc_params;

typedef struct {
  double h_old, h_new, dh, die_roll;
  int accept;
  double avg_plaq;         /* of the chain state after the accept step */
  double s_gauge;          /* gauge action of that state */
  double demon_e[2];       /* demon energies (plaquette, rectangle sums) */
  double reunit_dev;       /* max deviation from reunitarize */
  int status;              /* MAG_OK or MAG_ERR_NAN */
} mag_traj_result;

/* ---------------------------------------------------------------- context */
int mag_ctx_create(int device, mag_ctx* out);
int mag_ctx_destroy(mag_ctx ctx);
const char* mag_last_error(mag_ctx ctx);
int mag_ctx_synchronize(mag_ctx ctx);
/* Total device bytes held by a context's fields and workspace. */
int mag_ctx_memory(mag_ctx ctx, size_t* bytes);
/* Time of the kernels of the last timed region (see mag_ctx_set_timing):
 * per-kernel-family event time, milliseconds. */
int mag_ctx_set_timing(mag_ctx ctx, int on);
/* While timing is on, time only the launches of one kernel family (NULL or
 * "" = every family): fewer event records inside a timed region. */
int mag_ctx_set_timing_family(mag_ctx ctx, const char* family);
int mag_ctx_kernel_time(mag_ctx ctx, const char* family, double* ms, long* launches);
int mag_ctx_reset_timing(mag_ctx ctx);

/* ---------------------------------------------------------------- fields
 * qlat::Field<cps::Matrix>::init(geo) with multiplicity 4 (field-hmc.h:777-781) */
int mag_field_create(mag_ctx ctx, const int dims[4], int kind, int prec, mag_field* out);
int mag_field_destroy(mag_field f);
int mag_field_upload_aos(mag_field f, const double* host);
int mag_field_download_aos(mag_field f, double* host);
int mag_field_copy(mag_field dst, mag_field src); /* gFieldExt = gField (field-hmc.h:829) */
int mag_field_info(mag_field f, int dims[4], int* kind, int* prec, size_t* bytes);

/* ------------------------------------------------------------ link algebra */
int mag_unit_links(mag_field U);                   /* UnitMatrix start (quench_test/main.C:92-97) */
int mag_haar_links(mag_field U, uint64_t seed);    /* hot start; Philox Haar SU(3) */
/* init_momentum (field-hmc.h:646-693): P = sum_a T_a g_a, g ~ N(0, sqrt m) */
int mag_init_momentum(mag_field P, uint64_t seed, uint64_t traj, double m);
/* reunitarize (field-matrix.h:83-100) */
int mag_reunitarize(mag_field U, double* max_dev);
/* expiQ (field-hmc.h:82-123), batch of n 3x3 matrices [n][18]; prec selects
 * the arithmetic of the 3x3 algebra. Q is Hermitian traceless, as in the
 * reference: Q^2 is formed from its upper triangle and real diagonal. */
int mag_expiQ_batch(mag_ctx ctx, const double* Q, double* out, long n, int prec);
/* compute_Lambda (field-multi.h:180-269), batch of n; Q Hermitian traceless */
int mag_compute_Lambda_batch(mag_ctx ctx, const double* Q, const double* SigmaP, const double* U, double* out,
                             long n, int prec);
/* DexpiQ (field-multi.h:90-178), batch of n: d exp(iQ(t))/dt for dQ/dt = DQ,
 * zero when Q = 0 exactly (as the reference); Q Hermitian traceless */
int mag_DexpiQ_batch(mag_ctx ctx, const double* Q, const double* DQ, double* out, long n, int prec);

/* ------------------------------------------------------- forces and updates */
/* get_force (field-hmc.h:339-425); with n_stout > 0 the force of the smeared
 * action, by the stout chain of md::get_Fforce (field-multi-md.h:116-368) in
 * its isotropic form. F is MAG_ALG8 of U's precision. */
int mag_gauge_force(mag_field F, mag_field U, const mag_action* act);
/* evolve_gauge_field (field-hmc.h:442-482) */
int mag_evolve_gauge(mag_field U, mag_field P, double eps, int mag);
/* evolve_momentum (field-hmc.h:427-440) */
int mag_evolve_momentum(mag_field P, mag_field F, double eps, double m);
/* isotropic stout smearing, n levels; out has U's kind and precision */
int mag_stout_smear(mag_field out, mag_field U, double rho, int nlev);
/* force_gradient_integrator (field-hmc.h:525-567; md:: variant
 * field-multi-md.h:599-661); reunitarizes at the end when reunit != 0 */
int mag_fg_integrate(mag_field U, mag_field P, const mag_action* act, int steps, double dt, int reunit,
                     double* max_dev);
/* leap_frog_integrator (field-hmc.h:569-587) */
int mag_leapfrog_integrate(mag_field U, mag_field P, const mag_action* act, int steps, double dt);

/* -------------------------------------------------------------- reductions
 * All accumulate in FP64 on the device, in a fixed order. */
int mag_total_plaq(mag_field U, double* out);          /* field-matrix.h:299-330 */
int mag_total_rectangular(mag_field U, double* out);   /* field-matrix.h:354-385 */
int mag_avg_plaquette(mag_field U, double* out);       /* field-matrix.h:387-389 */
/* potential part of get_hamiltonian (field-hmc.h:632-640) on the smeared field */
int mag_gauge_action(mag_field U, const mag_action* act, double* S);
int mag_kinetic_energy(mag_field P, int mag, double* K);
/* get_hamiltonian (field-hmc.h:589-644): parts = {kinetic, potential} */
int mag_hamiltonian(mag_field U, mag_field P, const mag_action* act, double parts[2], double* H);
/* avg_wilson_loop (field-wloop.h:58-98) */
int mag_avg_wilson_loop(mag_field U, const int r[3], int T, double* out);
/* Asynchronous variants: the result stays on the device, written to d_out
 * (device memory) in the context's stream order; nothing synchronises.
 * total_plaq / total_rectangular / avg_wilson_loop: d_out[0];
 * hamiltonian: d_out[0..2] = {kinetic, potential, H}. */
int mag_total_plaq_async(mag_field U, double* d_out);
int mag_total_rectangular_async(mag_field U, double* d_out);
int mag_hamiltonian_async(mag_field U, mag_field P, const mag_action* act, double* d_out);
int mag_avg_wilson_loop_async(mag_field U, const int r[3], int T, double* d_out);

/* Allocation-free Hamiltonian. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). It computes what
 * mag_hamiltonian_async computes for a thin action (n_stout = 0), with the
 * same kernels in the same order (d_out[0..2] = {kinetic, potential, H}),
 * but its partial sums live in a fixed device array of the library (1.5 MiB
 * per GPU, held until the process ends) instead of context workspace, and it
 * allocates nothing. Lattices above MAG_NOALLOC_MAX_SITES sites return
 * MAG_ERR_ARG; smeared actions MAG_ERR_UNSUPPORTED. Every call on one GPU
 * shares that array: calls that overlap in time on one GPU (other streams,
 * contexts or threads) race, and the caller must serialise them. */
#define MAG_NOALLOC_MAX_SITES 16777216L /* 64^4 */
int mag_hamiltonian_noalloc(mag_field U, mag_field P, const mag_action* act, double* d_out);

/* -------------------------------------------------------------- HMC chain
 * run_chmc / update_field (field-hmc.h:770-986) as a device-resident chain:
 * links, backup, momenta, workspace and RNG state stay in HBM. */
int mag_chain_create(mag_ctx ctx, const int dims[4], const mag_action* act, const mag_hmc_params* p,
                     mag_chain* out);
int mag_chain_destroy(mag_chain c);
int mag_chain_set_links_aos(mag_chain c, const double* host);
int mag_chain_get_links_aos(mag_chain c, double* host);
int mag_chain_haar_links(mag_chain c, uint64_t seed);
/* Borrowed view of the chain's links. The demon window centre S0 and the
 * demon energies (field-demon.h:152-153) are taken from the links at the
 * first mag_chain_run or mag_chain_demon_state call and kept; set_links_aos
 * and haar_links drop them. After writing the links through this handle,
 * call mag_chain_reset_demon so the window is re-centred on the new links. */
int mag_chain_links(mag_chain c, mag_field* U);
int mag_chain_reset_demon(mag_chain c);
/* Run ntraj trajectories; results[i] for each (may be NULL). Trajectory
 * indices continue across calls (they key the momentum and die-roll streams). */
int mag_chain_run(mag_chain c, int ntraj, mag_traj_result* results);
/* FP64 reversibility check from the current state (its links widened to FP64
 * and projected onto SU(3)): trajectory forward, momenta flipped, trajectory
 * back. Returns max |U_fwd->bwd - U| and the Hamiltonian change of the round
 * trip; the chain state is unchanged. */
int mag_chain_reversibility(mag_chain c, double* max_link_dev, double* dh_roundtrip);
int mag_chain_demon_state(mag_chain c, double state[4]); /* S0p, S0r, Ep, Er */
int mag_chain_traj_index(mag_chain c, long* next_traj);

/* -------------------------------------------------------------- demon
 * demon_microcanonical (field-demon.h:25-126) in the reference's sequential
 * order (sites in index order, mu, 2 hits), executed by one wavefront on the
 * device; E[4] in/out (plaquette, rectangle, chair, twist). */
int mag_demon_microcanonical(mag_field U, int nsweeps, double E_max, uint64_t seed, double E[4],
                             long acc_rej[2]);

/* chair_staple_dagger / twist_staple_dagger (field-staple.h:141-165) for every
 * link: S(x, mu) = the sum over the 72 chair (twist) paths of field-staple.h:
 * 27-135, defects kept, returned undaggered as the reference returns it. U
 * and S are FP64 MAG_LINK18 fields of one shape (S holds general matrices). */
int mag_chair_staple_dagger(mag_field S, mag_field U);
int mag_twist_staple_dagger(mag_field S, mag_field U);

/* Checkerboard demon: demon_microcanonical as a parallel update with a demon
 * per link. Ep, Er, Ec, Et are MAG_REAL1 fields of U's shape holding every
 * link's plaquette, rectangle, chair and twist demon energies (in/out; start
 * them at zero, or at any values inside the window). Each hit proposes
 * random_su3_from_su2(0.9) U as the sequential demon does (same Philox
 * stream) and is accepted iff every demon of that link stays inside
 * (-E_max, E_max) after taking the action change. Per sweep and direction
 * the links go in 128 classes (x_mu mod 2, x_nu mod 4 for the transverse nu)
 * whose members share no staple path: the sum over links of S_k + E_k is
 * conserved for the closed loops (plaquette, rectangle) and the result does
 * not depend on thread order. Extents must be multiples of 4. acc_rej may be NULL. */
int mag_demon_checkerboard(mag_field U, mag_field Ep, mag_field Er, mag_field Ec, mag_field Et, int nsweeps,
                           double E_max, uint64_t seed, long acc_rej[2]);

/* -------------------------------------------------------- local Metropolis
 * algGCtrnHeatBath::run (alg_gchb.h:290-387) on the Wilson action, FP64
 * links: num_iter iterations of nHits Metropolis hits per link with
 * getRandomSU3BySubsu2(small) proposals (alg_gchb.h:32-96); a constrained
 * link (transverse coordinates multiples of mag) below the last one of its
 * coarse link is updated together with its successor (U1 eps, eps^+ U2).
 * mag = 0 updates every link alone. After each iteration: reunitarize and
 * record the average and constrained plaquettes (plaq, cons_plaq: num_iter
 * values each, may be NULL). acc_rej = {accept, reject, acceptCons,
 * rejectCons}. Order: the reference's sequential site order (one lane; the
 * same decisions as the oracle; at most 16^4 sites) or 16 checkerboard classes per iteration
 * (mu x parity of x_mu x parity of the transverse coordinate sum, whose
 * updates share no staple; even extents). Philox stream (seed, link, iteration). */
enum { MAG_ORDER_REFERENCE = 0, MAG_ORDER_CHECKERBOARD = 1 };
int mag_local_metropolis(mag_field U, double beta, int mag, int num_iter, int nhits, double small, uint64_t seed,
                         int order, long acc_rej[4], double* plaq, double* cons_plaq);

/* ------------------------------------------------------ mag-blocked lattices
 * Coarse/fine tools of the constrained HMC (link classes of is_constrained,
 * field-hmc.h:309-337). The fine field's extents are mag times the coarse
 * field's; both fields share format and context. */
/* coarse link = ordered product of the mag fine links from mag*xC
 * (coarse initialisation of double_multigrid, field-hmc.h:1068-1078) */
int mag_coarsen_links(mag_field fine, int mag, mag_field coarse);
/* update_constrain (field-hmc.h:988-1009): first fine link of each coarse
 * link := U_c [U(x+mu) ... U(x+(mag-1)mu)]^+ */
int mag_update_constrain(mag_field fine, mag_field coarse, int mag);
/* check_constrained_plaquette (field-matrix.h:407-443) */
int mag_constrained_plaquette(mag_field U, int mag, double* out);
/* lattice_expansion (expansion/main.C:19-62) without the file I/O: unit
 * links, the coarse link at fine site mag*x */
int mag_lattice_expansion(mag_field coarse, int mag, mag_field fine);

/* ------------------------------------------------ coupled coarse/fine HMC
 * The md:: HMC of field-multi-md.h, in FP64. Fine links UF (MAG_LINK18, even
 * extents) and coarse links UC of half the fine extents; coarse site c sits
 * at fine site 2c. Per coarse link: U2 = U(2c) U(2c+mu), Q = -i rho
 * TA(stp U2^+) with the six 2x1 staples stp (get_Q, field-multi.h:33-39),
 * Gb = e^{iQ} U2, S_xi = -(1 + md_alpha) sum xi Re Tr(UC^+ Gb). xi is a
 * MAG_REAL1 field of the coarse shape. The reference leaves rho and XI0 to
 * its includer and fixes md_alpha = 0.5 (field-multi-md.h:28). */
typedef struct {
  double rho;       /* stout weight of get_Q */
  double xi0;       /* XI0: the value init_xi writes, and the heat bath's weight */
  double md_alpha;  /* md_alpha */
} mag_multi_params;

/* get_xi_energy (field-multi-md.h:30-72) */
int mag_xi_energy(mag_field UF, mag_field UC, mag_field xi, const mag_multi_params* mp, double* out);
/* get_Q (field-multi.h:32-39) for every coarse link: Q = -i rho TA(stp U2^+)
 * into a coarse FP64 MAG_ALG8 field (Hermitian traceless, 8 components) */
int mag_multi_get_Q(mag_field Q, mag_field UF, const mag_multi_params* mp);
/* get_Gb (field-multi-md.h:771-783): Gb = e^{iQ} U2 into a coarse link field */
int mag_multi_gb(mag_field Gb, mag_field UF, const mag_multi_params* mp);
/* md::get_Fforce (field-multi-md.h:116-368): thin gauge force (no link
 * classes) plus the stout term of S_xi / (1 + md_alpha). F is fine MAG_ALG8;
 * act must be thin (n_stout = 0). */
int mag_multi_fforce(mag_field F, mag_field UF, mag_field UC, mag_field xi, const mag_action* act,
                     const mag_multi_params* mp);
/* md::get_Cforce (field-multi-md.h:398-425): coarse momentum
 * -i md_alpha/(1+md_alpha) (dGb/dt) Gb^+ along the fine momenta PF; CF is
 * coarse MAG_ALG8 */
int mag_multi_cforce(mag_field CF, mag_field UF, mag_field PF, const mag_multi_params* mp);
/* heatbath (field-multi-md.h:757-769): one simplest_metropolis hit per
 * coarse link with weight (1 + md_alpha) XI0 Re Tr(UC Gb^+); Philox stream
 * keyed by (seed, coarse link, traj) */
int mag_multi_heatbath(mag_field UC, mag_field UF, const mag_multi_params* mp, uint64_t seed, uint64_t traj);
/* force_gradient_integrator_zeta_T (field-multi-md.h:441-475); mag is
 * evolve_gauge_field's constraint block size (Arg_chmc::mag) */
int mag_multi_zeta_T(mag_field UF, mag_field PF, mag_field UC, double dt, int mag, const mag_multi_params* mp);
/* integrators: force_gradient_integrator_zeta_H (477-597), the md::
 * force_gradient_integrator with the coarse links fixed (599-661; dt is the
 * reference's Farg.dt) and nested_integrator (663-695). Both fields are
 * reunitarized at the end of each FG pass; max_dev = {fine, coarse}. */
enum { MAG_MULTI_ZETA_H = 0, MAG_MULTI_FG = 1, MAG_MULTI_NESTED = 2 };
int mag_multi_integrate(mag_field UF, mag_field PF, mag_field UC, mag_field xi, const mag_action* act,
                        const mag_multi_params* mp, int which, int steps, double dt, double max_dev[2]);
/* md::get_hamiltonian_multi (field-multi-md.h:74-114): parts = {kinetic
 * (all links), thin gauge action, S_xi} */
int mag_multi_hamiltonian(mag_field UF, mag_field PF, mag_field UC, mag_field xi, const mag_action* act,
                          const mag_multi_params* mp, double parts[3], double* H);

/* md::run_hmc_multi (field-multi-md.h:785-988) as a device-resident chain:
 * per trajectory init_momentum, heatbath, init_xi, H, zeta_H, H and the
 * Metropolis step on both fields (forced accepts first). A reject restores
 * both fields to the last accepted state, dropping that trajectory's heat-bath
 * update as the reference does. Uses trajectory_length, dt,
 * num_forced_accept and seed of mag_hmc_params; precision must be MAG_F64.
 * Both fields start as unit links; set them through mag_multi_fields. */
typedef struct {
  double h_old, h_new, dh, die_roll;
  int accept;
  double avg_plaq_fine, avg_plaq_coarse;  /* of the state after the accept step */
  double parts_old[3], parts_new[3];      /* kinetic, gauge, xi */
  double reunit_dev[2];                   /* fine, coarse */
  int status;                             /* MAG_OK or MAG_ERR_NAN */
} mag_multi_result;
int mag_multi_create(mag_ctx ctx, const int fine_dims[4], const mag_action* act, const mag_hmc_params* p,
                     const mag_multi_params* mp, mag_multi* out);
int mag_multi_destroy(mag_multi m);
int mag_multi_fields(mag_multi m, mag_field* UF, mag_field* UC); /* borrowed */
int mag_multi_run(mag_multi m, int ntraj, mag_multi_result* results);

/* ---------------------------------------------------------- NERSC config I/O
 * export_config_nersc / import_config_nersc (field-matrix.h:445-713): text
 * header BEGIN_HEADER ... END_HEADER, then the links in site order
 * (x fastest), mu 0..3 per site, row-major complex IEEE64 big-endian; 18
 * reals per link (DATATYPE 4D_SU3_GAUGE_3x3) or the first two rows, 12 reals
 * (4D_SU3_GAUGE, third row rebuilt by reunitarize on import). CHECKSUM is
 * the 32-bit sum of the data taken as 32-bit words of the host-order doubles
 * (what the reference's import checks, field-matrix.h:684, 697). The
 * reference's export of 4D_SU3_GAUGE sums the byte-swapped words instead
 * (field-matrix.h:473-477); the reader accepts that sum for 4D_SU3_GAUGE and
 * sets checksum_swapped. */
typedef struct {
  int dims[4];
  int skip_third;            /* 1: 4D_SU3_GAUGE (two rows); 0: 4D_SU3_GAUGE_3x3 */
  double link_trace;         /* LINK_TRACE: average Re Tr U / 3 */
  double plaquette;          /* PLAQUETTE: avg_plaquette */
  double beta;               /* BETA */
  int sequence_number;       /* SEQUENCE_NUMBER */
  char ensemble_label[128];  /* ENSEMBLE_LABEL */
  uint32_t checksum;         /* CHECKSUM */
  uint32_t crc32;            /* CRC32HASH: CRC-32 (IEEE) of the data in host (little-endian) byte order */
  int checksum_swapped;      /* read: 1 if CHECKSUM matched the sum over the byte-swapped words */
} mag_nersc_header;

/* Host-side file I/O of links in the qlat AoS layout [site][mu][3][3][re,im]
 * (no device work). Write fills h->checksum and h->crc32; read verifies the
 * checksum (aos may be NULL to read the header only; truncated files come
 * back with a zero third row). */
int mag_nersc_write(const char* path, const double* aos, mag_nersc_header* h);
int mag_nersc_read(const char* path, double* aos, mag_nersc_header* h);
/* export_config_nersc (field-matrix.h:452-568): plaquette and link trace of U
 * computed on the device, links written from U. */
int mag_export_config_nersc(mag_field U, const char* path, double beta, int sequence_number,
                            const char* ensemble_label, int skip_third);
/* import_config_nersc (field-matrix.h:592-713): dimensions must match U; the
 * checksum must match; truncated links are rebuilt by reunitarize; the
 * plaquette of U must agree with the header to 1e-4. h may be NULL. */
int mag_import_config_nersc(mag_field U, const char* path, mag_nersc_header* h);

/* -------------------------------------------------------------- serving
 * Continuous batching of independent chains (no counterpart in the
 * reference, whose statics allow one chain per process, field-hmc.h:911-922).
 * Chains are admitted in arrival order while their device footprint fits
 * the budget; chains of one shape class are batched into single launches
 * (optionally replayed from a HIP graph), and every batch runs its
 * trajectories on its own schedule on a pool of streams, so batches of
 * small lattices do not wait for batches of large ones. A chain's results
 * are the same alone or in any mix. */
typedef struct {
  int dims[4];
  mag_action act;
  mag_hmc_params hmc;
  int ntraj;             /* trajectories to run */
  uint64_t links_seed;   /* Haar-random start */
} mag_serve_request;

enum { MAG_CHAIN_PENDING = 0, MAG_CHAIN_RUNNING = 1, MAG_CHAIN_DONE = 2, MAG_CHAIN_FAILED = 3,
       MAG_CHAIN_CANCELLED = 4 };

int mag_server_footprint(const mag_serve_request* req, size_t* bytes);
int mag_server_create(mag_ctx ctx, size_t hbm_budget, int max_batch, int use_graphs, mag_server* out);
int mag_server_destroy(mag_server s);
int mag_server_submit(mag_server s, const mag_serve_request* req, int* id);
int mag_server_cancel(mag_server s, int id);
/* admit pending chains, launch every batch not on the device, wait until at
 * least one batch has finished its trajectory and process (accept, record,
 * retire) every batch that has */
int mag_server_step(mag_server s, int* active, int* pending);
int mag_server_status(mag_server s, int id, int* state, int* done_traj);
/* per-trajectory results and completion times (seconds since server creation) */
int mag_server_results(mag_server s, int id, mag_traj_result* out, double* t_done, int max);
int mag_server_info(mag_server s, size_t* used_bytes, long* trajectories, double* elapsed_s);
/* host seconds spent in admission, batch builds (recording, table upload,
 * graph capture), launches, waiting for completions and accept/retire
 * processing; builds = number of batch (re)builds */
int mag_server_host_times(mag_server s, double seconds[5], long* builds);

#ifdef __cplusplus
}
#endif
#endif /* MAG_AMD_H */
