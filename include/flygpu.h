/*
 * flygpu.h -- C-ABI of libflygpu, the MI355X (gfx950) implementation of openFLY's saddle-search hot path:
 * periodic cell-list neighbour lists, EAM/Finnis-Sinclair energy, gradient and Hessian-vector products,
 * and the L-BFGS / dimer steps that drive them.
 *
 * Conventions
 *  - Every function returns a status code (FLYG_OK on success).  The message of the last failure on a
 *    context is returned by flyg_ctx_last_error().  No C++ exception crosses this boundary.
 *  - Per-atom vectors are interleaved xyz, 3*n doubles, exactly the layout of the reference's SoA
 *    properties Position / Delta / PotentialGradient / Axis (property.hpp:53-85, adaptor.hpp:46-109).
 *  - Arguments tagged `where` are host pointers when where == FLYG_HOST and device pointers (on the
 *    context's device, ordered on the context's stream) when where == FLYG_DEVICE.
 *  - The box basis is a 3x3 row-major matrix whose COLUMNS are the basis vectors (the reference's
 *    Box(Mat basis, Arr<bool> periodic), box.hpp:103-120).
 *  - Handles are single-owner and not thread-safe, like the reference objects they replace; use one
 *    context per host thread (the reference's per-thread ThreadData, find.hpp:310-317).
 */
#ifndef FLYGPU_H
#define FLYGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define FLYG_OK 0
#define FLYG_INVALID_ARG 1  /* reference: fly::RuntimeError thrown by verify()/error() (core.hpp:89-138) */
#define FLYG_SPLINE_OOB 2   /* reference: fly::Spline::OutOfBounds (spline.hpp:46-51, thrown at :141-143) */
#define FLYG_DEVICE_ERROR 3 /* HIP runtime failure */
#define FLYG_BUSY 4         /* serving: the device memory budget cannot take this now; retry after searches finish */
#define FLYG_STOPPED 5      /* a step observer (flyg_*_set_observer) asked the run to stop */

/* Pointer location for `where` arguments. */
#define FLYG_HOST 0
#define FLYG_DEVICE 1

/* Pair arithmetic precision of a potential handle. */
#define FLYG_FP64 0  /* plain form: everything in FP64 */
#define FLYG_MIXED 1 /* FP32 per-pair spline/force arithmetic, FP64 accumulation and FP64 energies */

/* Dimer exit codes.  The first five are the reference's Dimer::Exit (dimer.hpp:123-129) with the sign
 * dropped (reference value = -code: success 0, convex -1, iter_max -2, collision -3, uninit -4), so a code
 * can index a histogram; the C++ facade's Dimer::Exit carries the reference's own values.  The last two
 * occur only in batched searches. */
#define FLYG_EXIT_SUCCESS 0
#define FLYG_EXIT_CONVEX 1
#define FLYG_EXIT_ITER_MAX 2
#define FLYG_EXIT_COLLISION 3
#define FLYG_EXIT_UNINIT 4
#define FLYG_EXIT_REGION 5 /* left the rows of its active region; the library re-runs it with larger skins, so
                              a finished search reports it only if the region outgrew half the box */
#define FLYG_EXIT_OOB 6    /* a spline was evaluated outside its table (Spline::OutOfBounds, caught by
                              Master::find_batch, find.cpp:718-723) after the allowed re-perturbations */

typedef struct flyg_ctx_s* flyg_ctx;
typedef struct flyg_eam_s* flyg_eam;
typedef struct flyg_list_s* flyg_list;
typedef struct flyg_steplbfgs_s* flyg_steplbfgs;
typedef struct flyg_lbfgs_s* flyg_lbfgs;
typedef struct flyg_dimer_s* flyg_dimer;
typedef struct flyg_rotor_s* flyg_rotor;
typedef struct flyg_batch_s* flyg_batch;
typedef struct flyg_server_s* flyg_server;

/* ---------------------------------------------------------------------------------------------------
 * Context: a device, one HIP stream (default priority) and the last error message.
 * No reference counterpart; stands in for the per-thread state of Master (find.hpp:310-317).
 * ------------------------------------------------------------------------------------------------- */
int flyg_ctx_create(int device, flyg_ctx* out);
int flyg_ctx_destroy(flyg_ctx ctx);
const char* flyg_ctx_last_error(flyg_ctx ctx);
int flyg_ctx_synchronize(flyg_ctx ctx);
/* The hipStream_t of the context, for callers that time or order work around it. */
void* flyg_ctx_stream(flyg_ctx ctx);
/* Device name and compute-unit count. */
int flyg_ctx_device_info(flyg_ctx ctx, char* name, size_t len, int* compute_units);

/* ---------------------------------------------------------------------------------------------------
 * Potential tables: replaces potential::DataEAM(Options, std::ifstream) (data.hpp:54, data.cpp:67-153)
 * and the EAM(TypeMap, shared_ptr<DataEAM>) ctor (eam.hpp:50-63).  Parses a setfl/eam.fs file on the
 * host (3 comment lines, species line, "Nrho drho Nr dr rcut", per-species "Z mass" + F + rho tables,
 * then r*V pair tables for j <= i, five values per line), builds the natural cubic splines and
 * uploads FP64 and FP32 coefficient tables.
 * ------------------------------------------------------------------------------------------------- */
int flyg_eam_load(flyg_ctx ctx, const char* path, flyg_eam* out);
/* The same from the file's text (for DataEAM(Options, std::ifstream)). */
int flyg_eam_load_text(flyg_ctx ctx, const char* text, size_t len, flyg_eam* out);
int flyg_eam_destroy(flyg_eam eam);
/* potential::EAM::r_cut() (eam.hpp:70) / DataEAM::type_map(): names is ntypes*16 chars (NUL padded). */
int flyg_eam_info(flyg_eam eam, int* ntypes, double* r_cut, double* masses, char* names);
/* FLYG_FP64 (default) or FLYG_MIXED for gradient / HVP pair arithmetic.  Energies are always FP64. */
int flyg_eam_set_precision(flyg_eam eam, int precision);
/* Lanes of a wavefront that share one atom's neighbour row in the gradient kernels: 4, 8 or 16, or 0
 * (default) to time the three once per system size.  Results are bitwise identical for every choice. */
int flyg_eam_set_lanes(flyg_eam eam, int lanes);
/* Kernel timing of gradient calls (no reference counterpart; for benchmarks): while on, each gradient
 * call records a HIP event pair on the context's stream around its density + force kernels.
 * set_timing(1) also clears the record; flyg_eam_timing synchronises the stream and returns the summed
 * event time (ms) and the number of timed calls since. */
int flyg_eam_set_timing(flyg_eam eam, int on);
int flyg_eam_timing(flyg_eam eam, double* ms_total, int64_t* calls);
/* Evaluates one table on the device at m host points: kind 0 = F_a, 1 = phi(a,b), 2 = V(a,b);
 * deriv 0/1/2 = f/f'/f'' (DataEAM::f/phi/v + Spline::f/fp/fpp, data.hpp:73-116, spline.hpp:81-123). */
int flyg_eam_table_eval(flyg_eam eam, int kind, int a, int b, const double* x, int64_t m, int deriv, double* out);

/* ---------------------------------------------------------------------------------------------------
 * Neighbour list: replaces neigh::List (list.hpp:71-210, list.cpp:37-165).
 * ------------------------------------------------------------------------------------------------- */
/* List(Box const&, double r_cut) (list.hpp:86-91). */
int flyg_list_create(flyg_ctx ctx, const double basis[9], const int periodic[3], double r_cut, flyg_list* out);
int flyg_list_destroy(flyg_list nl);
/* List::rebuild(SoA<Position const&>, int threads) (list.hpp:99, list.cpp:53-95). */
int flyg_list_rebuild(flyg_list nl, const double* pos, int64_t n, int where);
/* List::update(SoA<Delta const&>) (list.hpp:119, list.cpp:37-51): r <- r - delta for every atom. */
int flyg_list_update(flyg_list nl, const double* delta, int64_t n, int where);
/* List::size() (list.hpp:161). */
int flyg_list_size(flyg_list nl, int64_t* n);
/* Number of ghost images made by the last rebuild (List::m_num_plus_ghosts - size(), list.hpp:178). */
int flyg_list_num_ghosts(flyg_list nl, int64_t* n_ghosts);
/* Pairs with r < r_cut (List::for_neighbours, list.hpp:138-154): count, then export as CSR rows with
 * the real neighbour index, the image shift code (s0+1)+3(s1+1)+9(s2+1) and dr = r_n - r_i. */
int flyg_list_count(flyg_list nl, double r_cut, int64_t* n_pairs);
int flyg_list_export(flyg_list nl, double r_cut, int64_t* offsets, int32_t* j, int32_t* shift, double* dr);
/* Current internal positions (canonical images at the last rebuild, minus later updates). */
int flyg_list_positions(flyg_list nl, double* pos, int where);

/* ---------------------------------------------------------------------------------------------------
 * EAM evaluation: replaces potential::EAM / potential::Generic energy, gradient, hessian
 * (generic.hpp:79-136, eam.cpp:33-233).  `types` uint32 and `frozen` uint8 per atom.  Passing
 * types = NULL and frozen = NULL (both) to any call of a potential handle that takes them selects the
 * copy staged in the handle by flyg_eam_stage_atoms (or by the last call with host arrays).
 * ------------------------------------------------------------------------------------------------- */
/* No counterpart in the reference, whose SoA inputs are host arrays read on every call: copies the species
 * and frozen flags of n atoms (host or device memory, `where`) into the potential handle's device memory.
 * Later calls that pass types = frozen = NULL use this copy; the gradient then gathers it into a list's
 * cell-order layout once per list build instead of on every call, as it does for host arrays (device
 * arrays passed to each call may change between calls, so they are gathered every time).  Host inputs
 * are validated here (FLYG_INVALID_ARG); device inputs by the kernels that read them.  A device source
 * is copied asynchronously on the context's stream: work on other streams must not write it until that
 * stream has passed the copy. */
int flyg_eam_stage_atoms(flyg_eam eam, const uint32_t* types, const uint8_t* frozen, int64_t n, int where);
/* EAM::energy (eam.cpp:33-51): sum over all atoms, frozen included. */
int flyg_eam_energy(flyg_eam eam, flyg_list nl, const uint32_t* types, const uint8_t* frozen, int64_t n, int where,
                    double* energy);
/* EAM::gradient (eam.cpp:53-120): two passes (density then force); frozen atoms get exactly 0. */
int flyg_eam_gradient(flyg_eam eam, flyg_list nl, const uint32_t* types, const uint8_t* frozen, int64_t n, int where,
                      double* grad);
/* No counterpart in the reference: EAM::gradient for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same kernels and results as
 * flyg_eam_gradient with FLYG_DEVICE pointers (types, frozen, grad on the context's device), but the
 * density pass's output goes to a library-held device array of at most 131072 atoms (32 B per atom, one
 * copy per GPU, held until the process ends) instead of a per-handle buffer, and the lanes per atom are
 * not tuned (set_lanes, an earlier tuning for this size, else 8).  Larger systems are refused
 * (FLYG_INVALID_ARG).  Calls that overlap on one GPU share that array and race.  Asynchronous on the
 * context's stream; spline and type errors surface at the next synchronising call. */
int flyg_eam_gradient_noalloc(flyg_eam eam, flyg_list nl, const uint32_t* types, const uint8_t* frozen, int64_t n,
                              double* grad);
/* H v with H the analytic Hessian of EAM::hessian (eam.cpp:122-233) restricted to non-frozen atoms,
 * computed matrix-free in two passes (no 3N x 3N matrix). */
int flyg_eam_hvp(flyg_eam eam, flyg_list nl, const uint32_t* types, const uint8_t* frozen, int64_t n,
                 const double* v, double* hv, int where);

/* EAM::hessian (eam.cpp:122-233) of the non-frozen atoms, mass-weighted with the potential's masses
 * (Hessian::mass_weight, hessian.hpp:136-143) and diagonalised (Hessian::eigenvalues, hessian.hpp:86-93):
 * the 3M eigenvalues in increasing order (M = non-frozen atoms; *n_dof = 3M, eigenvalues holds 3n).
 * These are the reference's 3n eigenvalues without the 3 per frozen atom that its zero rows add.  The
 * block is assembled on the device from 3M Hessian-vector products and diagonalised there by cyclic
 * Jacobi rotations; at most 16384 degrees of freedom. */
int flyg_eam_hessian_eigenvalues(flyg_eam eam, flyg_list nl, const uint32_t* types, const uint8_t* frozen, int64_t n,
                                 int where, double* eigenvalues, int64_t* n_dof);
/* Master's reading of a spectrum (find.cpp:849-866 at a saddle, 1113-1126 at a minimum): the sum of
 * log(lambda) over eigenvalues above zero_tol (Master::Options::hessian_eigen_zero_tol, 1e-6), and the
 * numbers of eigenvalues within zero_tol of zero and below -zero_tol. */
int flyg_log_prod_eigen(const double* eigenvalues, int64_t n, double zero_tol, double* log_prod, int64_t* n_zero,
                        int64_t* n_negative);
/* Mechanism::kinetic_pre (find.cpp:287): sqrt(exp(log_prod_min - log_prod_sp) / (2 pi amu)), amu in kg. */
int flyg_prefactor(double log_prod_min, double log_prod_sp, double* prefactor);

/* ---------------------------------------------------------------------------------------------------
 * StepLBFGS: replaces minimise::StepLBFGS (lbfgs_core.hpp:37-119, lbfgs_core.cpp:29-87).
 * ------------------------------------------------------------------------------------------------- */
int flyg_steplbfgs_create(flyg_ctx ctx, int m, flyg_steplbfgs* out);
int flyg_steplbfgs_destroy(flyg_steplbfgs s);
int flyg_steplbfgs_clear(flyg_steplbfgs s);
/* newton_step(x, g): returns q ~ H^-1 g (the caller moves x <- x - q). x, g, q hold len doubles. */
int flyg_steplbfgs_newton_step(flyg_steplbfgs s, const double* x, const double* g, int64_t len, double* q, int where);

/* ---------------------------------------------------------------------------------------------------
 * LBFGS minimiser: replaces minimise::LBFGS (lbfgs.hpp:53-118, lbfgs.cpp:29-115).
 * ------------------------------------------------------------------------------------------------- */
typedef struct {
  int n;               /* history length (10) */
  int iter_max;        /* (2000) */
  double f2norm;       /* (1e-6) */
  double skin_frac;    /* (1.1) */
  double proj_tol;     /* (0) */
  double max_trust;    /* (0.5) */
  double min_trust;    /* (0.01) */
  double grow_trust;   /* (1.5) */
  double shrink_trust; /* (0.5) */
} flyg_lbfgs_options;

void flyg_lbfgs_default_options(flyg_lbfgs_options* opt);
/* LBFGS(Options, Box) */
int flyg_lbfgs_create(flyg_ctx ctx, const flyg_lbfgs_options* opt, const double basis[9], const int periodic[3],
                      flyg_lbfgs* out);
int flyg_lbfgs_destroy(flyg_lbfgs m);
/* LBFGS::minimise(out, in, pot, threads) -> bool: *failed = 1 when iter_max was reached (the
 * reference returns true on failure, lbfgs.cpp:114).  pos is in/out, grad out. */
int flyg_lbfgs_minimise(flyg_lbfgs m, flyg_eam eam, double* pos, double* grad, const uint32_t* types,
                        const uint8_t* frozen, int64_t n, int where, int* failed, int* iterations);
/* Step observer, the reference's Options::fout trajectory dump (LBFGS: lbfgs.cpp:75-77; Dimer:
 * dimer.cpp:88-90): called on the calling thread at the start of every iteration of flyg_lbfgs_minimise /
 * flyg_dimer_find_sp, before the convergence test, with the iteration number and the current positions
 * in host memory (3n doubles, valid during the call).  A run that converges at iteration k calls it k + 1
 * times.  A nonzero return stops the run, which then returns FLYG_STOPPED.  Costs one device-to-host copy
 * of the positions per iteration; fn = NULL removes the observer. */
typedef int (*flyg_step_observer)(void* user, int iteration, const double* xyz, int64_t n);
int flyg_lbfgs_set_observer(flyg_lbfgs m, flyg_step_observer fn, void* user);

/* ---------------------------------------------------------------------------------------------------
 * Dimer saddle search: replaces saddle::Rotor + saddle::Dimer (rotor.hpp:42-139, dimer.hpp:51-170,
 * rotor.cpp:49-164, dimer.cpp:28-171).
 * ------------------------------------------------------------------------------------------------- */
typedef struct {
  int n;               /* L-BFGS history (10) */
  int max_steps;       /* (1000) */
  int hist_check_freq; /* (5) */
  double f2norm;       /* (1e-5) */
  double skin_frac;    /* (1.1) */
  double proj_tol;     /* (0) */
  double max_trust;    /* (0.5) */
  double min_trust;    /* (0.05) */
  double grow_trust;   /* (1.5) */
  double shrink_trust; /* (0.5) */
  int convex_max;      /* (3) */
  int use_history;     /* (1) */
} flyg_dimer_options;

typedef struct {
  int n;               /* rotation L-BFGS history (6) */
  int iter_max_rot;    /* (20) */
  double delta_r;      /* dimer half-length (1e-3) */
  double theta_tol;    /* (pi/360) */
  int relax_in_convex; /* (1) */
} flyg_rotor_options;

typedef struct {
  int exit_code;   /* FLYG_EXIT_* */
  int steps;       /* translation steps taken */
  int64_t grads;   /* gradient evaluations */
  int rebuilds;    /* neighbour-list rebuilds */
  double curv;     /* last curvature along the axis */
  double energy;   /* batched searches: FP64 barrier E(saddle) - E(base); plain dimer: 0 */
  double max_disp; /* batched searches: largest displacement of a mobile atom from its base position */
  int launches;    /* batched searches: kernel launches used (a search that outgrew its region runs again in a
                      larger one, continuing from where it left: see resumes) */
  int rot_iters;   /* batched searches: rotor iterations */
  double device_ms; /* batched searches: device time its workgroups spent on the search (all launches since its
                       start state) */
  double device_grad_ms; /* ... of which in gradient evaluations */
  int retries;     /* batched searches started by Master::perturb: re-perturbations after a spline error */
  int resumes;     /* batched searches: launches that continued from the state saved at a region exit (the rest
                      of launches - 1 restarted from the start state) */
} flyg_search_stats;

void flyg_dimer_default_options(flyg_dimer_options* opt);
void flyg_rotor_default_options(flyg_rotor_options* opt);
/* Rotor(Options) (rotor.hpp:67). */
int flyg_rotor_create(flyg_ctx ctx, const flyg_rotor_options* opt, flyg_rotor* out);
int flyg_rotor_destroy(flyg_rotor r);
/* Rotor::r_cut(pot) (rotor.hpp:123): pot.r_cut() + 2 delta_r, the least cut-off of the list eff_gradient uses. */
int flyg_rotor_r_cut(flyg_rotor r, flyg_eam eam, double* r_cut);
/* Rotor::eff_gradient(out, inout, in, pot, nl, count_frozen, threads) -> curvature (rotor.hpp:105-111,
 * rotor.cpp:49-164): the list must be ready (rebuilt or updated) at the positions, with a cut-off of at
 * least flyg_rotor_r_cut; it is moved by +-delta_r along the axis and returned to the same state, never
 * rebuilt.  axis (3n) is rotated towards the lowest curvature mode in place; eff (3n) receives
 * g - 2 (g.n) n (or -(g.n) n at positive curvature without relax_in_convex), with its atom-mean removed
 * when count_frozen == 0.  *curv: the curvature along the returned axis. */
int flyg_rotor_eff_gradient(flyg_rotor r, flyg_eam eam, flyg_list nl, const uint32_t* types, const uint8_t* frozen,
                            int64_t n, double* axis, double* eff, int count_frozen, int where, double* curv);
/* Dimer(Options, Rotor::Options, Box) */
int flyg_dimer_create(flyg_ctx ctx, const flyg_dimer_options* opt, const flyg_rotor_options* ropt,
                      const double basis[9], const int periodic[3], flyg_dimer* out);
int flyg_dimer_destroy(flyg_dimer d);
/* Dimer::find_sp(out, in, in_min, pot, hist_sp, theta_tol, count_frozen, threads) -> Exit.
 * pos/axis are in/out (3n each); in_min 3n; hist_sp n_hist * 3n (may be NULL when n_hist == 0). */
int flyg_dimer_find_sp(flyg_dimer d, flyg_eam eam, double* pos, double* axis, const uint32_t* types,
                       const uint8_t* frozen, int64_t n, const double* in_min, const double* hist_sp, int n_hist,
                       double theta_tol, int count_frozen, int where, flyg_search_stats* stats);
/* Step observer of flyg_dimer_find_sp (see flyg_lbfgs_set_observer): the dimer's centre positions. */
int flyg_dimer_set_observer(flyg_dimer d, flyg_step_observer fn, void* user);

/* ---------------------------------------------------------------------------------------------------
 * Batched saddle searches: the device-side counterpart of Master::find_batch's task-per-search loop
 * (find.cpp:684-755, one OpenMP task per search, each running Master::find_sp -> Dimer::find_sp with
 * threads = 1, find.cpp:983-1009).  One persistent kernel advances many searches at once, one
 * workgroup per search, each on its active region (mobile atoms + frozen shells) held on the device.
 * ------------------------------------------------------------------------------------------------- */
typedef struct {
  int precision;        /* FLYG_MIXED (default) or FLYG_FP64 */
  int block_threads;    /* workgroup size, 0 = autotuned */
  int lanes_per_atom;   /* lanes cooperating on one atom's neighbours, 0 = autotuned */
  double list_skin;     /* Verlet skin of the per-search local lists (A), 0 = default */
  int max_concurrent;   /* resident searches per launch, 0 = 2 per compute unit */
} flyg_batch_options;

void flyg_batch_default_options(flyg_batch_options* opt);
/* One base system (the relaxed minimum min0), resident on the device for the handle's lifetime. */
int flyg_batch_create(flyg_ctx ctx, flyg_eam eam, const flyg_dimer_options* opt, const flyg_rotor_options* ropt,
                      const flyg_batch_options* bopt, const double basis[9], const int periodic[3],
                      const double* min_pos, const uint32_t* types, int64_t n, flyg_batch* out);
int flyg_batch_destroy(flyg_batch b);
/* Queue searches.  For search k: every atom farther than radius[k] from atom centre[k] (minimum
 * image, in min_pos) is frozen; start_pos / start_axis hold the full 3n start state of that search
 * (a perturbed copy of min_pos and its dimer axis, as Master::perturb writes them). */
int flyg_batch_submit(flyg_batch b, int count, const int64_t* centre, const double* radius, const double* start_pos,
                      const double* start_axis, int64_t* ids);
/* flyg_batch_submit with Dimer::find_sp's saddle-history test (dimer.cpp:102-131): every hist_check_freq
 * steps a search whose direction from the minimum (x - min_pos) is within theta_tol of the direction to
 * one of the n_hist known saddles (hist_pos: n_hist * 3n full positions, shared by these searches) stops
 * with FLYG_EXIT_COLLISION.  theta_tol = 0 or n_hist = 0: no test. */
int flyg_batch_submit_ex(flyg_batch b, int count, const int64_t* centre, const double* radius, const double* start_pos,
                         const double* start_axis, double theta_tol, const double* hist_pos, int n_hist, int64_t* ids);
/* Master::find_batch's searches (find.cpp:684-732): search k starts from flyg_perturb of min_pos around
 * centre[k] (atoms beyond radius[k] frozen) drawn from its own Xoshiro256** state prng[4k .. 4k+3], which
 * advances and is written back; a search that evaluates a spline outside its table is re-perturbed from
 * the same stream and run again, at most three times (find.cpp:706-723; stats.retries counts them, and
 * FLYG_EXIT_OOB reports the fourth failure, which the reference throws).  History as flyg_batch_submit_ex. */
int flyg_batch_submit_perturbed(flyg_batch b, int count, const int64_t* centre, const double* radius, uint64_t* prng,
                                double stddev, double r_pert, double theta_tol, const double* hist_pos, int n_hist,
                                int64_t* ids);
/* Run every queued search to completion on the device. */
int flyg_batch_run(flyg_batch b);
/* Results of search `id`: exit code / stats (energy = FP64 E(sp) - E(min0), the barrier), and the full
 * 3n saddle positions and axis (either may be NULL). */
int flyg_batch_result(flyg_batch b, int64_t id, flyg_search_stats* stats, double* pos, double* axis);
/* Mobile / region atom counts of search `id`. */
int flyg_batch_region(flyg_batch b, int64_t id, int* n_mobile, int* n_region);
/* Workgroup shape the batch ran with (0 before its first run): set in its options, else taken from an
 * earlier batch over the same potential, system size and precision, else autotuned on the first run.
 * No reference counterpart. */
int flyg_batch_shape(flyg_batch b, int* block_threads, int* lanes_per_atom);

/* ---------------------------------------------------------------------------------------------------
 * Mechanisms: Master::do_adj_min + the pathway checks of Master::saddle_2_mech (find.cpp:757-837,
 * 1014-1090) on the device, right after a search converges: relax from the saddle nudged by
 * +-nudge_frac |sp - min| along the axis (LBFGS::minimise with `minimiser`), call the minimum nearer
 * the initial one rev, remove the centroid drift of rev (over all atoms), and require
 * |rev - min| <= basin_tol, |fwd - rev| >= basin_tol, |rev - sp| and |fwd - sp| >= stationary_tol, the
 * atom moving most from rev to fwd to be the search's centre, and E(sp) > E(rev).
 * ------------------------------------------------------------------------------------------------- */
typedef struct {
  int enabled;                   /* 0 (default): no mechanism step */
  double nudge_frac;             /* Master::Options::nudge_frac (find.hpp:184), 0.01 */
  double basin_tol;              /* Master::Options::basin_tol (find.hpp:112), 0.05 A */
  double stationary_tol;         /* Master::Options::stationary_tol (find.hpp:120), basin_tol / 2 */
  flyg_lbfgs_options minimiser;  /* LBFGS::Options of the two relaxations (history length: the dimer's).
                                    The relaxations use the batch's precision: FP32 pair forces carry
                                    rounding noise near 1e-5 eV/A, so FLYG_MIXED requires f2norm >= 1e-4
                                    (the reference default 1e-6 is for FP64) and is refused otherwise. */
} flyg_mech_options;

typedef struct {
  int code;          /* 0 pathway; 1 a minimisation failed; 2 unlinked (rev away from the initial minimum);
                        3 converged back to the initial minimum; 4 min->sp or 5 sp->min displacement below
                        stationary_tol; 6 centred on another atom; 7 E(sp) <= E(rev); -1 not checked */
  int min_iters;     /* L-BFGS iterations of the two relaxations */
  double barrier;    /* E(sp) - E(rev), FP64 */
  double delta;      /* E(fwd) - E(rev), FP64 */
  double drift[3];   /* centroid drift of rev from the initial minimum; the pathway is rev, sp, fwd - drift */
} flyg_mechanism;

void flyg_mech_default_options(flyg_mech_options* opt);
int flyg_batch_set_mechanisms(flyg_batch b, const flyg_mech_options* opt);
/* Mechanism of search `id` (stats.exit_code 0 only); rev_pos / fwd_pos: full 3n positions of the two
 * minima before the drift correction (either may be NULL). */
int flyg_batch_mechanism(flyg_batch b, int64_t id, flyg_mechanism* mech, double* rev_pos, double* fwd_pos);

/* ---------------------------------------------------------------------------------------------------
 * Serving: continuous batching of saddle-search requests from many systems (a kMC loop's catalogue
 * misses, Master::find_mechs -> find_batch, find.cpp:211-312,684-755, one system per request stream).
 * Each admitted system stays resident on the device (positions, base densities, types) until removed
 * and idle; each search keeps its region lists and L-BFGS histories on the device until it ends.
 * Queued searches of any systems are packed into launches of the persistent search kernel on up to
 * `streams` streams; a stream that finishes takes whatever is queued then, so requests that arrive
 * while others run join the next launch and their work overlaps on the device.  Admission is gated
 * by a device memory budget.  Results are bitwise independent of what else runs beside a search.
 * No reference counterpart as an API: the reference runs the same searches as OpenMP tasks.
 * ------------------------------------------------------------------------------------------------- */
typedef struct {
  int precision;        /* FLYG_MIXED (default) or FLYG_FP64 */
  int block_threads;    /* workgroup size, 0 = autotuned on the first launch */
  int lanes_per_atom;   /* lanes per atom row, 0 = autotuned */
  double list_skin;     /* initial Verlet skin of the per-search lists (A), 0 = default */
  int max_concurrent;   /* resident workgroups per launch, 0 = 2 per compute unit */
  int streams;          /* concurrent launches, 0 = 4 (launches overlap only on streams that have a hardware
                           queue each: HIP gives a process GPU_MAX_HW_QUEUES of them, 4 unless set) */
  int max_batch;        /* searches per launch, 0 = 4 per compute unit */
  int64_t mem_budget;   /* device bytes for systems + launches, 0 = 80% of the free memory at creation */
} flyg_server_options;

typedef struct {
  int64_t queued, in_flight, systems, device_bytes, budget_bytes, launches, searches_done;
  double kernel_ms;     /* summed kernel time of finished launches */
  int busy_streams, block_threads, lanes_per_atom;
  double host_ms;       /* host time spent building search regions and packing launches */
} flyg_server_state;

void flyg_server_default_options(flyg_server_options* opt);
int flyg_server_create(flyg_ctx ctx, flyg_eam eam, const flyg_dimer_options* opt, const flyg_rotor_options* ropt,
                       const flyg_server_options* sopt, flyg_server* out);
int flyg_server_destroy(flyg_server s);
/* Admit a system (a relaxed minimum); FLYG_BUSY when the memory budget cannot hold it now. */
int flyg_server_add_system(flyg_server s, const double basis[9], const int periodic[3], const double* min_pos,
                           const uint32_t* types, int64_t n, int64_t* sys_id);
/* The system is freed once none of its searches is queued or running. */
int flyg_server_remove_system(flyg_server s, int64_t sys_id);
/* Queue `count` searches on a system.  Search k freezes every atom farther than radius[k] from
 * centre[k]; its start state is the base positions except the atoms kick_atoms[kick_offsets[k] ..
 * kick_offsets[k+1]) at kick_pos (3 per atom), and its dimer axis is kick_axis on those atoms, zero
 * elsewhere (the sparse form of what Master::perturb writes, find.cpp:1147-1214). */
int flyg_server_submit(flyg_server s, int64_t sys_id, int count, const int64_t* centre, const double* radius,
                       const int64_t* kick_offsets, const int64_t* kick_atoms, const double* kick_pos,
                       const double* kick_axis, int64_t* ids);
/* Queue `count` searches on a system started by Master::perturb, as flyg_batch_submit_perturbed: search k
 * draws from its own Xoshiro256** state prng[4k .. 4k+3] (advanced and written back) and is re-perturbed
 * after a spline error, at most three times. */
int flyg_server_submit_perturbed(flyg_server s, int64_t sys_id, int count, const int64_t* centre, const double* radius,
                                 uint64_t* prng, double stddev, double r_pert, int64_t* ids);
/* Harvest finished launches and start queued work on free streams; waits up to wait_ms for at least one
 * search to finish.  Writes up to `cap` finished search ids. */
int flyg_server_progress(flyg_server s, double wait_ms, int64_t* done_ids, int cap, int* n_done);
/* Result of a finished search: stats (energy = FP64 barrier), mobile atom count, and (any may be NULL)
 * the mobile atoms' global indices, saddle positions and axis (3 per mobile atom). */
int flyg_server_result(flyg_server s, int64_t id, flyg_search_stats* stats, int* n_mobile, int64_t* atoms,
                       double* pos, double* axis);
/* Forget a finished search's result. */
int flyg_server_release(flyg_server s, int64_t id);
int flyg_server_set_mechanisms(flyg_server s, const flyg_mech_options* opt);
/* Mechanism of a finished search; rev / fwd: 3 per mobile atom (the order of flyg_server_result's atoms). */
int flyg_server_mechanism(flyg_server s, int64_t id, flyg_mechanism* mech, double* rev, double* fwd);
int flyg_server_status(flyg_server s, flyg_server_state* state);

/* ---------------------------------------------------------------------------------------------------
 * Search starts: Master::perturb (find.cpp:1147-1214) and find_batch's history tolerance (find.cpp:692).
 * ------------------------------------------------------------------------------------------------- */
/* Master::perturb of `pos` around atom `centre` (host arrays, orthorhombic box): sd ~ N(stddev, stddev/3)
 * redrawn until positive; each non-frozen atom at r < r_pert from the centre (every periodic image, in
 * ascending (index, r) order) moves by N(0, sd) (1 - r/r_pert) per component and gets N(0, 1) axis
 * components; the centre moves by N(0, sd) and gets N(0, 1) axis components; the axis is normalised.
 * Gaussians come from the Marsaglia polar method on Xoshiro256** (random.hpp:96-161) with the state
 * prng[4], which advances (a later call continues the stream, as Master's per-thread generator does).
 * out_pos / out_axis: 3n each; *sd (may be NULL) receives the sampled standard deviation. */
int flyg_perturb(const double basis[9], const int periodic[3], const double* pos, const uint8_t* frozen, int64_t n,
                 int64_t centre, double stddev, double r_pert, uint64_t prng[4], double* out_pos, double* out_axis,
                 double* sd);
/* Master::find_batch's collision angle after `tot` searches at one site: ((30 - 2.5) e^(-0.02 tot) + 2.5)
 * degrees, in radians (find.cpp:692). */
int flyg_theta_tol(int tot, double* theta_tol);

#ifdef __cplusplus
}
#endif

#endif /* FLYGPU_H */
