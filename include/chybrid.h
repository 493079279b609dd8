/*
 * chybrid.h -- C ABI of the MI355X (gfx950) implementation of the
 * collisional-hybrid per-timestep hot path.
 *
 * One opaque context (chy_ctx) owns every device allocation of one run.
 * Host pointers passed in are borrowed for the duration of the call; outputs
 * are caller-allocated.  Every entry point returns 0 on success or a negative
 * error code whose magnitude is the PETSc error class the reference raises in
 * the same situation (reference hybrid.h / PetscErrorCode convention);
 * chy_last_error() returns the message.  No exception or abort crosses the
 * ABI.  A context is used by one host thread at a time; calls are ordered on
 * the context's HIP stream and are complete when they return.
 *
 * Layout contract (reference setup.c:6-146, hybrid.h:40-45,140-143):
 *   grid fields   [Nz][Ny][Nx] f64, DMDA natural order (x fastest);
 *   vlasov        [Nz][Ny][Nx][4] f64 interleaved {n, Gx, Gy, Gz} (GridNode);
 *   particles     RealVector AoS [np][4] f64 {x, y, z, r} for position and
 *                 velocity (DMSwarm fields "position" and "velocity"; r is
 *                 written as the magnitude and ignored on input).
 */
#ifndef CHYBRID_H
#define CHYBRID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (negated PETSc classes) -------------------------------- */
#define CHY_OK                 0
#define CHY_ERR_MEM          (-55)  /* PETSC_ERR_MEM */
#define CHY_ERR_SUP          (-56)  /* PETSC_ERR_SUP */
#define CHY_ERR_SIG          (-59)  /* PETSC_ERR_SIG: collision refusals, particles.c:603-605 */
#define CHY_ERR_ARG_WRONG    (-62)  /* PETSC_ERR_ARG_WRONG: parameters.c:454,468,505 */
#define CHY_ERR_ARG_RANGE    (-63)  /* PETSC_ERR_ARG_OUTOFRANGE */
#define CHY_ERR_FP           (-72)  /* PETSC_ERR_FP: NaN/Inf detected on the GPU */
#define CHY_ERR_ARG_INCOMP   (-75)  /* PETSC_ERR_ARG_INCOMP: parameters.c:486,493,500 */
#define CHY_ERR_LIB          (-76)  /* PETSC_ERR_LIB: HIP runtime failure */
#define CHY_ERR_NOT_CONVERGED (-91) /* PETSC_ERR_NOT_CONVERGED: -ksp_error_if_not_converged */

/* ---- enumerations (reference lhs.h:8-12, rhs.h:8-12, hybrid.h:84-102) ---- */
enum { CHY_LHS_IDENTITY = 0, CHY_LHS_LAPLACIAN = 1, CHY_LHS_FULL = 2 };
enum { CHY_RHS_CONSTANT = 0, CHY_RHS_SINUSOIDAL = 1, CHY_RHS_FULL = 2 };
/* --density-type (parameters.c:14-16); CHY_DENSITY_LATTICE is a build
 * addition: ppc ions per cell at seeded uniform offsets, selected by -ppc */
enum { CHY_DENSITY_FLAT_SOBOL = 0, CHY_DENSITY_FLAT_REVERSE, CHY_DENSITY_FLAT_NORMAL,
       CHY_DENSITY_UNIFORM, CHY_DENSITY_SINUSOIDAL, CHY_DENSITY_GAUSSIAN, CHY_DENSITY_LATTICE };
enum { CHY_BC_PERIODIC = 0, CHY_BC_INJECTION, CHY_BC_REFLECTION, CHY_BC_DIRICHLET, CHY_BC_NEUMANN };
/* build additions */
enum { CHY_PREC_FP64 = 0, CHY_PREC_MIXED = 1 };
enum { CHY_KSP_GMRES = 0, CHY_KSP_BCGS = 1 };
enum { CHY_PC_NONE = 0, CHY_PC_JACOBI = 1, CHY_PC_FFT = 2, CHY_PC_ZLINE = 3, CHY_PC_CHEBY = 4, CHY_PC_FFT2 = 5,
       CHY_PC_FFT2Z = 6, CHY_PC_MG = 7 };
/* -ksp_gmres_cgs_refinement_type: REFINE_NEVER, REFINE_IFNEEDED (when the
 * first pass removes more than half of the vector's squared norm), REFINE_ALWAYS */
enum { CHY_CGS_REFINE_NEVER = 0, CHY_CGS_REFINE_IFNEEDED = 1, CHY_CGS_REFINE_ALWAYS = 2 };
/* -ksp_initial_guess: the previous step's phi (default), zero (the reference's
 * ComputeInitialPhi, potential.c:20-29), or the linear extrapolation
 * 2 phi_n - phi_(n-1) of the last two steps */
enum { CHY_GUESS_PREVIOUS = 0, CHY_GUESS_ZERO = 1, CHY_GUESS_EXTRAPOLATE = 2 };
enum { CHY_PERT_NONE = 0, CHY_PERT_SINUSOIDAL_X = 1, CHY_PERT_UNIFORM = 2, CHY_PERT_GAUSSIAN = 3 };

/* ---- run parameters: the scalars of the reference Context (hybrid.h:106-138)
 * plus their derived values (parameters.c:507-556, setup.c:43-49). ---------- */
typedef struct {
  /* Grid (hybrid.h:47-53) */
  int64_t N[3];          /* nodes per axis */
  double  p0[3], p1[3];  /* lower / upper bounds [m] */
  double  L[3], d[3];    /* derived: p1 - p0, L / N */
  int32_t bc[3][2];      /* CHY_BC_*; only all-periodic is supported */
  /* electrons (Charged) */
  double  qe, me, nue, Te, gamma_e;
  double  ve0[3], veT[3], veT_r;
  double  Omega_e[3], kappa_e[3];
  /* ions (Charged) */
  double  qi, mi, nui, Ti;
  double  vi0[3], viT[3], viT_r;
  double  Omega_i[3], kappa_i[3];
  /* neutrals (Neutral, hybrid.h:67-72) */
  double  mn, Tn, vnT;
  double  vn0[3], vn0_r;
  /* plasma (Plasma, hybrid.h:74-78) */
  double  B0[3], E0[3];
  int64_t Np;
  /* time loop (hybrid.h:130-133) */
  int64_t Nt;
  double  dt;
  /* operator / forcing / initial density selection */
  int32_t lhs_type, rhs_type, density_type;
  int32_t view_lhs;
  /* solver (PETSc KSP/PC subset mapped onto the GPU solver) */
  int32_t ksp_type, pc_type;
  double  ksp_rtol, ksp_atol;
  int64_t ksp_max_it, ksp_restart;
  int32_t ksp_monitor, ksp_converged_reason, ksp_error_if_not_converged;
  int32_t pc_cheby_degree; /* Chebyshev-Jacobi preconditioner: polynomial degree */
  double  ksp_inner_rtol;  /* mixed precision: floor of the FP32 inner relative tolerance */
  int32_t ksp_gmres_cgs_refinement; /* CHY_CGS_REFINE_*: second Gram-Schmidt pass (PETSc KSPGMRESCGSRefinementType) */
  int32_t ksp_initial_guess; /* CHY_GUESS_*: initial phi of each step's solve */
  /* build additions: seed, particles per cell, perturbation, precision, output cadence */
  uint64_t seed;
  int64_t ppc;
  int32_t perturbation;
  double  perturbation_amplitude;
  int32_t weighted;      /* allocate per-particle weights (implied by a perturbation) */
  int32_t precision;
  int64_t output_interval;
  int64_t collision_refusal_limit; /* refused scatterings in one step that abort it (particles.c:603: 10); 0 = never */
  /* solver binary (parameters.c:419-440) */
  double  flux_scale[3];
  char    input[4096];
  char    output[4096];
  /* build addition: chy_step fills its diagnostics every this many steps (0 = never) */
  int64_t diagnostics_interval;
} chy_params;

typedef struct chy_ctx chy_ctx;

typedef struct {
  int64_t accepted;      /* collisions applied this step */
  int64_t refused;       /* refused (final speed > 10 vT), particles.c:597-605 */
  int64_t capped;        /* Bernoulli probabilities clipped at 1 */
  int64_t Nc;            /* target count, particles.c:514 */
  double  expected;      /* sum over the ions of their collision probability p: the
                            mean of accepted + refused (Nc up to the step's drift of S) */
  double  variance;      /* sum of p (1 - p): its variance */
} chy_collision_stats;

typedef struct {
  int64_t iterations;    /* total inner Krylov iterations */
  int64_t outer;         /* FP64 defect-correction sweeps (1 in FP64 mode) */
  double  residual;      /* true ||b - A phi|| / ||b|| in FP64 */
  int32_t converged;     /* 1 converged, 0 not (max_it reached) */
  int32_t reason;        /* PETSc-style KSPConvergedReason (>0 ok, <0 diverged) */
  double  ms;            /* wall time of the solve (RHS + Krylov + projections) */
} chy_solve_stats;

/* Global diagnostics, summed in FP64 on the device (chy_diagnostics). */
typedef struct {
  int64_t ions;            /* live ions */
  int64_t step;            /* collision step counter when taken */
  double  weight;          /* sum of the ion weights */
  double  momentum[3];     /* m_i sum w v  [kg m/s] */
  double  kinetic_energy;  /* 1/2 m_i sum w |v|^2  [J] */
  double  field_energy;    /* 1/2 eps0 sum |grad phi|^2 dV over the nodes  [J] */
  int32_t valid;           /* 1 when filled */
} chy_diag_stats;

typedef struct {
  chy_collision_stats collisions;
  chy_solve_stats     solve;
  double              ms_particles;  /* push + sort + deposit */
  double              ms_total;
  chy_diag_stats      diag;          /* after the step, when -diagnostics_interval selects it */
} chy_step_stats;

/* ---- parameters ----------------------------------------------------------
 * Defaults of ProcessOptions (reference parameters.c:28-441). */
int chy_params_default(chy_params *prm);
/* Parse an options-database style argument vector (reference flag names of
 * parameters.c plus -ksp_* / -pc_* and the build additions), then derive
 * (parameters.c:442-556, setup.c:9-49).  Unknown options are ignored like
 * PETSc's, except that err gets a warning list.  Replaces ProcessOptions. */
int chy_params_from_options(int argc, const char *const *argv, chy_params *prm,
                            char *err, size_t errlen);
/* Recompute derived values after editing a chy_params by hand. */
int chy_params_derive(chy_params *prm, char *err, size_t errlen);
/* Write the options.txt report (reference parameters.c:566-654). */
int chy_write_options_txt(const chy_params *prm, int argc, const char *const *argv, const char *path);

/* ---- context --------------------------------------------------------------
 * Replaces SetUpVlasovDM / SetUpIonsDM / SetUpPotentialDM / KSPCreate
 * (reference hybrid.c:52-78, setup.c:6-146).  np_capacity = 0 uses prm->Np. */
int chy_create(const chy_params *prm, int64_t np_capacity, chy_ctx **ctx);
int chy_destroy(chy_ctx *ctx);
const char *chy_last_error(const chy_ctx *ctx);
int chy_sync(chy_ctx *ctx);
/* The context's HIP stream (a hipStream_t), for callers that enqueue work. */
void *chy_stream(chy_ctx *ctx);
int chy_get_params(const chy_ctx *ctx, chy_params *prm);
int64_t chy_num_particles(const chy_ctx *ctx);
/* Step counter used to key the collision random streams. */
int chy_set_step(chy_ctx *ctx, int64_t it);

/* ---- particles (DMSwarm fields "position", "velocity"; setup.c:130-132) --- */
/* pos, vel: [np][4] f64; w may be NULL (unit weights); pid may be NULL
 * (pid = array index).  Particles are re-sorted by cell on upload. */
int chy_set_particles(chy_ctx *ctx, int64_t np, const double *pos, const double *vel,
                      const double *w, const uint32_t *pid);
/* Any output may be NULL.  Order is the current cell-sorted order. */
int chy_get_particles(chy_ctx *ctx, double *pos, double *vel, double *w, uint32_t *pid);
/* Seeded placement on the device, replacing the once-per-run
 * InitializePositions / InitializeVelocities (particles.c:9-132), by
 * prm->density_type:
 *   flat-sobol (default)  Np ions at the points of the reference's Sobol
 *                         sequence, v = r (L + d) - d/2 pulled inside the
 *                         domain (distributions.c:95-141; at most 2^30 - 1);
 *   uniform               Np / nc ions at the centre of every cell except the
 *                         last along each axis (distributions.c:33-92): the
 *                         ion count becomes Np/nc (Nx-1)(Ny-1)(Nz-1);
 *   sinusoidal            rejection sampling of 1 + 0.25 sin(2 pi x)
 *                         (distributions.c:144-232), against its maximum;
 *   lattice (-ppc P)      P ions per cell at uniform offsets (build addition);
 *   flat-normal, flat-reverse, gaussian: CHY_ERR_ARG_WRONG "Not implemented",
 *                         as particles.c:28-45.
 * Velocities vi0 + viT N(0,1) and the weight perturbation come from Philox
 * streams keyed by the seed (the reference's gasdev stream is not
 * reproduced); weights from prm perturbation. */
int chy_init_particles(chy_ctx *ctx);

/* ---- the hot path --------------------------------------------------------
 * UpdateVelocities (particles.c:628-643): gather E = E0 - grad(phi) at each
 * ion, Boris rotation, then collisions when mn > 0. */
int chy_update_velocities(chy_ctx *ctx, chy_collision_stats *stats);
/* UpdatePositions (particles.c:647-718): x += v dt, periodic wrap, then the
 * stable counting sort by cell that replaces DMSwarmMigrate. */
int chy_update_positions(chy_ctx *ctx);
/* CollectVlasovQuantities (particles.c:136-255): deterministic CIC deposit of
 * density and flux.  vlasov may be NULL; otherwise [Nz][Ny][Nx][4] f64. */
int chy_collect_moments(chy_ctx *ctx, double *vlasov);
/* Upload moments (solver input, fileio.c:6-70) / download the current ones. */
int chy_set_moments(chy_ctx *ctx, const double *vlasov);
int chy_get_moments(chy_ctx *ctx, double *vlasov);
/* ComputeRHS -> rhsFunc (potential.c:33-42, rhs.c): b [Nz][Ny][Nx] may be NULL. */
int chy_compute_rhs(chy_ctx *ctx, double *b);
/* ComputeLHS -> lhsFunc (potential.c:46-55, lhs.c) as a matrix-free product
 * with the operator of the current density: y = A(n) x, host arrays. */
int chy_lhs_apply(chy_ctx *ctx, const double *x, double *y);
int chy_lhs_apply_transpose(chy_ctx *ctx, const double *x, double *y);
/* ComputePotential (potential.c:6-16): RHS from the current moments, then the
 * Krylov solve warm-started from the previous phi (zero on the first call,
 * potential.c:20-29).  phi may be NULL. */
int chy_solve_potential(chy_ctx *ctx, double *phi, chy_solve_stats *stats);
/* Solve A(n) phi = b for a given b; phi in/out (the initial guess).  For the
 * full and Laplacian operators, whose null space is the constants, the mean
 * of b is removed first, as PETSc's KSP projects the right-hand side of a
 * matrix with an attached null space (lhs.c:334-335,530-531). */
int chy_solve(chy_ctx *ctx, const double *b, double *phi, chy_solve_stats *stats);
/* KSPSetTolerances (the KSP that potential.c:6-16 solves with): relative and
 * absolute tolerance and iteration limit of the following solves; a negative
 * value keeps the current setting (PETSC_DEFAULT). */
int chy_set_tolerances(chy_ctx *ctx, double rtol, double atol, int64_t max_it);
int chy_get_potential(chy_ctx *ctx, double *phi);
int chy_set_potential(chy_ctx *ctx, const double *phi);
int chy_get_rhs(chy_ctx *ctx, double *b);
/* PCApply of the selected preconditioner (the apply callback of a PETSc
 * PCShell, for a host that keeps its own KSP): z = M^-1 r for the current
 * moments, [Nz][Ny][Nx] FP64 in and out (the mean of z is arbitrary: the
 * operator annihilates constants).  none: z = r; jacobi: D^-1 r; fft:
 * A(mean n)^+ r; zline: cyclic tridiagonal solves of every z-line block of
 * A(n); cheby: pc_cheby_degree steps of Chebyshev iteration on D^-1 A from
 * zero, eigenvalues in [0.2, 2.2]; fft2: fft, then 2 steps of flexible GMRES
 * on the z-averaged residual with the z-averaged operator (the x-y coupling,
 * which the Hall term makes density dependent); fft2z (default): fft2, then the
 * zline solve of its residual added; mg (-pc_type gamg, hypre and ml map
 * here): one V(2,2) cycle of x-y semicoarsening multigrid with the damped
 * z-line smoother, any axis lengths. */
int chy_pc_apply(chy_ctx *ctx, const double *r, double *z);
/* One time step of the loop body hybrid.c:108-117 (velocities, positions,
 * moments, potential), fused on the device. */
int chy_step(chy_ctx *ctx, chy_step_stats *stats);

/* Global diagnostics of the current state (no reference counterpart; the
 * reference reports collision counts only, particles.c:525,620): ion count,
 * total weight, momentum and kinetic energy of the ions, and the energy of
 * the field of the current potential (periodic central differences, the node
 * field the push gathers, without the imposed uniform E0), each summed in
 * FP64 over a fixed grid, so repeated calls agree bitwise. */
int chy_diagnostics(chy_ctx *ctx, chy_diag_stats *d);
/* The same diagnostics without any allocation or host synchronisation, for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops).  No reference counterpart.  Enqueued on the context stream; writes
 * d_out[0..5] = {weight, momentum x, y, z, kinetic energy, field energy}, the
 * values chy_diagnostics returns bit for bit, to caller device memory.  The
 * block partials live in a static 48 KiB device array, one per GPU, set up
 * by chy_create; its reduction grid is fixed at 1024 blocks, so every ion and
 * node count fits.  Each call overwrites every partial it reads, so nothing
 * needs clearing, but calls that overlap on one GPU (two contexts' streams
 * at once) race on that array. */
int chy_diagnostics_noalloc(chy_ctx *ctx, double *d_out);

/* ---- adjoint of the deterministic step map --------------------------------
 * No reference counterpart: the reference has no derivatives.  The forward
 * map starts from the current cell-sorted ions: CIC moments and the potential
 * solve from a zero guess (hybrid.c:94-99), then nsteps steps of the loop
 * body (hybrid.c:101-123).  The objective is the final wave energy
 * J = sum_c (n_c - mean n)^2 of the deposited density.
 * chy_adjoint_forward runs the map, keeps a checkpoint every ckpt steps and
 * returns J.  chy_adjoint_backward returns dJ/dE0 (3 values) and, when
 * dJ_ddelta is not NULL, dJ/d delta_c for every cell ([Nz][Ny][Nx]), where
 * delta_c is an additive perturbation of the weight of each ion in cell c at
 * the start of the forward run.  Collision events and scattering angles are
 * frozen (common random numbers).  Segments between checkpoints are
 * recomputed; the context is left at the end state of the forward run.
 * FP64 contexts with the full operator and forcing only (CHY_ERR_SUP). */
int chy_adjoint_forward(chy_ctx *ctx, int64_t nsteps, int64_t ckpt, double *J);
int chy_adjoint_backward(chy_ctx *ctx, double dJ_dE0[3], double *dJ_ddelta);

/* ---- output (OutputHDF5, reference fileio.c:73-126) -------------------------
 * Writes <stem><insert>.hdf with datasets "density", "x flux", "y flux",
 * "z flux", "rhs" and "potential", each (Nz, Ny, Nx) f64.  HDF5 is loaded at
 * run time; without it the six arrays go to <stem><insert>.<name>.f64 raw
 * files and the call returns CHY_ERR_SUP after writing them. */
int chy_output_write(chy_ctx *ctx, const char *stem, const char *insert);
/* The same files written behind the computation (OutputHDF5 at scale): the
 * six fields are snapshotted on the device in stream order, so the following
 * calls may change them at once; the copy to pinned host memory runs on a
 * second stream and a host thread writes the file.  Returns without waiting.
 * One output is in flight at a time: the next call (or chy_output_write)
 * first waits for the previous one and returns its error, if any.  Holds
 * 6 nc f64 of device and of pinned host memory from the first call on. */
int chy_output_write_async(chy_ctx *ctx, const char *stem, const char *insert);
/* Wait for the asynchronous output in flight, if any; its result (CHY_ERR_SUP
 * after writing raw files without HDF5, as chy_output_write). */
int chy_output_wait(chy_ctx *ctx);

/* ---- input (LoadVlasovQuantities, reference fileio.c:6-70) ------------------
 * Reads the density dataset "density-kji" (nc = Nx Ny Nz f64 values, x
 * fastest) of the HDF5 file at path into vlasov [Nz][Ny][Nx][4] as
 * {n, s_x n, s_y n, s_z n}, s = prm->flux_scale (fluxes made from the density,
 * fileio.c:44-53).  A path ending in ".f64" is read instead as nc raw f64
 * values.  HDF5 is loaded at run time; without it an HDF5 path is
 * CHY_ERR_SUP.  A missing file or dataset, or one whose size is not nc, is
 * CHY_ERR_ARG_WRONG; msg (may be NULL) receives a description of any failure. */
int chy_read_vlasov(const chy_params *prm, const char *path, double *vlasov, char *msg, size_t msg_len);

/* ---- instrumentation -------------------------------------------------------
 * Per-kernel HIP-event timing on the context stream (off by default).
 * Names: "efield", "push", "sort_scan", "sort_scatter", "sort_reorder",
 * "deposit_cells", "deposit_nodes", "rhs", "pc", "lhs_apply", "krylov".
 * Events are resolved when read, so timing adds no synchronisation. */
int chy_timing_enable(chy_ctx *ctx, int on);
/* Restrict timing to the comma-separated timer names (NULL: all), so a timed
 * run can keep event records on one kernel only. */
int chy_timing_select(chy_ctx *ctx, const char *names);
int chy_timing_get(chy_ctx *ctx, const char *name, double *total_ms, int64_t *count);
int chy_timing_reset(chy_ctx *ctx);
/* Plain device-to-device copies of nbytes on the context stream, repeat
 * times each: a grid-stride vector kernel over 16384 blocks ("copy" timer)
 * and one 16-byte element per thread ("copy1" timer, the streaming
 * ceiling); used to calibrate HBM byte counters and rooflines. */
int chy_copy_kernel(chy_ctx *ctx, int64_t nbytes, int repeat);

#ifdef __cplusplus
}
#endif
#endif
