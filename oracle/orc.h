/*
 * orc.h -- CPU oracle for the collisional-hybrid per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64 restatement of the
 * reference algorithm (myoung-space-science/collisional-hybrid, src/),
 * written from its formulas, used solely as the checker: by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  The
 * product library (libchybrid.so) never links, calls or falls back to it.
 *
 * The reference itself cannot be compiled here (every file includes
 * <petsc.h>; PETSc/SLEPc are absent), so each function below cites the
 * reference file:line whose mathematics it restates and is pinned by the
 * closed forms, invariants and brute-force checks in tests/test_oracle_*.py.
 *
 * Semantics are the reference's *intended* ones where its code is defective
 * (undefined behaviour in BorisMover's v-minus loop, the zero cross product,
 * positions never stored, moments accumulated across steps, E0 missing from
 * the push, the d2Fdyz stencil typo).  Each such choice is marked "intended".
 *
 * Conventions shared with the GPU library:
 *  - periodic node-centred grid, N[a] nodes per axis, node (i,j,k) at
 *    p0 + (i,j,k)*d, linear index c = i + Nx*(j + Ny*k) (DMDA natural order);
 *  - cell (i,j,k) spans nodes i..i+1 etc. and shares the linear index of its
 *    lower corner node;
 *  - particle arrays are SoA, FP64; optional weight array w (NULL = 1);
 *  - all sums run in the fixed orders documented at each function, so the
 *    FP64 GPU kernels (built with -ffp-contract=off) can match them exactly.
 */
#ifndef CHY_ORACLE_H
#define CHY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- physical constants, reference constants.h:5-8 ---------------------- */
#define ORC_Q  1.6022e-19
#define ORC_MP 1.6726e-27
#define ORC_ME 9.1094e-31
#define ORC_KB 1.3807e-23

/* ---- counter-based RNG (Philox4x32-10, Salmon et al., SC'11) ------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* 24-bit uniform in (0,1): ((w>>8)+0.5)*2^-24, exact in FP32 and FP64. */
double orc_u24(uint32_t w);
double orc_normaliser(int64_t np, const double *vx, const double *vy, const double *vz, const double vn0[3],
                      double vnT);

/* ---- operator and forcing (reference lhs.c, rhs.c, calculus.c) ---------- */
enum { ORC_LHS_IDENTITY = 0, ORC_LHS_LAPLACIAN = 1, ORC_LHS_FULL = 2 };
enum { ORC_RHS_CONSTANT = 0, ORC_RHS_SINUSOIDAL = 1, ORC_RHS_FULL = 2 };

/* The 19 coefficients of row c of the full operator, in the column order of
 * lhs.c:11-104 (x+,x-,y+,y-,z+,z-, xy edges ++,+-,-+,--, xz edges, yz edges,
 * centre).  Restates lhs.c:405-511. */
void orc_lhs_full_row(const int64_t N[3], const double d[3], const double K[3],
                      const double *n, int64_t c, double coef[19]);
/* Offsets (di,dj,dk) of the 19 columns above. */
void orc_stencil_offsets(int off[19][3]);
/* y = A x for lhs type 0/1/2 (identity, laplacian lhs.c:205-218, full). */
void orc_lhs_apply(int type, const int64_t N[3], const double d[3], const double K[3],
                   const double *n, const double *x, double *y);
/* y = A^T x (matrix-free, from the rows' coefficients). */
void orc_lhs_apply_transpose(int type, const int64_t N[3], const double d[3], const double K[3],
                             const double *n, const double *x, double *y);
/* Dense assembly, row-major Nc x Nc (small grids only; N[a] >= 3). */
void orc_lhs_dense(int type, const int64_t N[3], const double d[3], const double K[3],
                   const double *n, double *A);
/* Forcing b from moments; rhs.c:8-303.  cth = gamma_e kB Te / Q,
 * cG = me nue / Q.  Full and sinusoidal forcings remove the mean (rhs.c:294-296,
 * 120-123); the constant one does not (rhs.c:33-38). */
void orc_rhs(int type, const int64_t N[3], const double d[3], const double K[3],
             const double E0[3], double cth, double cG,
             const double *n, const double *gx, const double *gy, const double *gz,
             double *b);
void orc_remove_mean(int64_t nc, double *v);

/* ---- particles (reference particles.c, vectors.c) ------------------------ */
typedef struct {
  int64_t N[3];
  double  p0[3], p1[3], d[3], L[3];
  double  qm;          /* ion charge / mass */
  double  dt;
  double  B0[3], E0[3];
  /* collisions (reference particles.c:471-624); enabled when mn > 0 */
  int     collide;
  double  mi, mn;
  double  viT;         /* ion thermal speed vT.r, reference parameters.c:541-545 */
  double  vnT;         /* neutral thermal speed, parameters.c:553 */
  double  vn0[3];
  int64_t Nc;          /* collisions per step, particles.c:510-519 */
  uint64_t seed;
  uint32_t step;
} orc_push_params;

/* Cell coordinates of position x along one axis: integer cell and fraction. */
void orc_cell_coord(double x, double p0, double d, int64_t N, int64_t *i, double *f);
/* Periodic wrap into [p0, p1) (intended form of particles.c:679-704). */
double orc_wrap(double x, double p0, double p1, double L);
/* Node electric field E = E0 - grad(phi), central differences, periodic. */
void orc_efield_nodes(const int64_t N[3], const double d[3], const double E0[3],
                      const double *phi, double *ex, double *ey, double *ez);
/* Trilinear gather of node fields at one particle, corner order q=0..7. */
void orc_gather(const int64_t N[3], const double p0[3], const double d[3],
                const double *ex, const double *ey, const double *ez,
                double x, double y, double z, double E[3]);
/* Closed-form mean relative speed E|v - v_n|, v_n ~ N(vn0, vnT^2 I). */
double orc_mean_rel_speed(const double v[3], const double vn0[3], double vnT);
/* Collision count per step, particles.c:510-519. */
int64_t orc_collision_count(int64_t Np, double nui, double dt, double viT,
                            const double vi0[3], const double vn0[3], double vnT,
                            double mi, double mn);

/* One push of every particle: gather E at the current position, Boris
 * rotation, collisions, move, new cell key.  S = sum of mean relative speeds
 * over the velocities entering this step (the collision normaliser).
 * Outputs: updated arrays in place, key[np], *S_next (sum over the final
 * velocities), stats[3] = {accepted, refused, capped}. */
void orc_push(const orc_push_params *P, const double *ex, const double *ey, const double *ez,
              int64_t np, double *x, double *y, double *z,
              double *vx, double *vy, double *vz, const uint32_t *pid,
              double S, int32_t *key, double *S_next, int64_t stats[3]);
/* Stable counting sort by key: perm[slot] = old index; counts per cell. */
void orc_sort_perm(int64_t ncell, int64_t np, const int32_t *key, int64_t *perm, int64_t *offsets);
/* Apply a permutation to a double / uint32 array (out[s] = in[perm[s]]). */
void orc_permute_d(int64_t np, const int64_t *perm, const double *in, double *out);
void orc_permute_u(int64_t np, const int64_t *perm, const uint32_t *in, uint32_t *out);
/* CIC deposit of particles sorted by cell (offsets[ncell+1]).  Per-cell
 * corner partial sums in array order, then per-node sums over corners
 * q = 0..7.  Restates particles.c:181-242 with fresh moments (intended). */
void orc_deposit_sorted(const int64_t N[3], const double p0[3], const double d[3],
                        int64_t np, const double *x, const double *y, const double *z,
                        const double *vx, const double *vy, const double *vz, const double *w,
                        const int64_t *offsets,
                        double *n, double *gx, double *gy, double *gz);
/* Same moments by direct scatter in particle order (the reference's loop
 * structure); used to cross-check the sorted deposit. */
void orc_deposit_scatter(const int64_t N[3], const double p0[3], const double d[3],
                         int64_t np, const double *x, const double *y, const double *z,
                         const double *vx, const double *vy, const double *vz, const double *w,
                         double *n, double *gx, double *gy, double *gz);

/* Seeded initial state: ppc particles per cell (pid = ppc*cell + slot),
 * offsets U[0,1)^3, velocity v0 + vT*N(0,1) per component; weights from the
 * perturbation (0 none, 1 sinusoidal-x, 2 uniform per cell, 3 gaussian per
 * cell) with amplitude amp.  w may be NULL when pert == 0. */
void orc_init_particles(const int64_t N[3], const double p0[3], const double d[3],
                        int64_t ppc, uint64_t seed, const double v0[3], const double vT[3],
                        int pert, double amp,
                        double *x, double *y, double *z, double *vx, double *vy, double *vz,
                        uint32_t *pid, double *w);

/* The reference's initial densities (particles.c:26-45, distributions.c):
 * flat-sobol (its Sobol sequence, distributions.c:95-141), uniform (cell
 * centres, distributions.c:33-92) and sinusoidal (rejection against
 * 1 + 0.25 sin(2 pi x), distributions.c:144-232).  Velocities and weights as
 * orc_init_particles.  Returns the ion count placed, -1 when the Sobol
 * sequence runs out of bits, -2 for another kind. */
enum { ORC_DENSITY_FLAT_SOBOL = 0, ORC_DENSITY_UNIFORM = 3, ORC_DENSITY_SINUSOIDAL = 4 };
void orc_sobol_dirs(uint32_t V[30][3]);
int orc_sobol_points(int64_t np, double *r);
int64_t orc_density_count(int kind, const int64_t N[3], int64_t Np);
int64_t orc_init_density(int kind, const int64_t N[3], const double p0[3], const double p1[3], const double L[3],
                         const double d[3], int64_t Np, uint64_t seed, const double v0[3], const double vT[3],
                         int pert, double amp, double *x, double *y, double *z, double *vx, double *vy,
                         double *vz, uint32_t *pid, double *w);

/* ---- potential solve (reference potential.c:6-16 + PETSc KSP) ----------- */
enum { ORC_PC_NONE = 0, ORC_PC_JACOBI = 1, ORC_PC_FFT = 2, ORC_PC_ZLINE = 3, ORC_PC_CHEBY = 4, ORC_PC_FFT2 = 5,
       ORC_PC_FFT2Z = 6, ORC_PC_MG = 7 };
/* multigrid (ORC_PC_MG): levels, smoothing sweeps before and after the
 * coarse correction, sweeps on the coarsest level, damping */
#define ORC_MG_MAX_LEVELS 8
#define ORC_MG_SWEEPS 2
#define ORC_MG_COARSE_SWEEPS 8
#define ORC_MG_OMEGA 0.8
int orc_mg_levels(const int64_t N[3], int64_t out[ORC_MG_MAX_LEVELS][3]);
void orc_mg_restrict(const int64_t Nf[3], const double *f, double *c);
void orc_mg_prolong_add(const int64_t Nf[3], const double *c, double *f);
enum { ORC_KSP_GMRES = 0, ORC_KSP_BCGS = 1 };
#define ORC_CHEBY_DEGREE 4
#define ORC_FFT2_ITS 2
/* Apply the preconditioner: z = M^{-1} r (mean removed).
 *   jacobi  z = D^-1 r
 *   fft     z = A(mean n)^+ r by 3-D FFT
 *   zline   block Jacobi over z-lines: the cyclic tridiagonal system of each
 *           (i, j) column (diagonal and +-z entries of A(n)) solved exactly
 *   cheby   `degree` steps of Chebyshev iteration on D^-1 A z = D^-1 r from
 *           z = 0, for eigenvalues in [0.2, 2.2] (0.1 and 1.1 times the
 *           Gershgorin bound 2 of D^-1 A)
 *   fft2    fft, then a correction of the z-averaged (kz = 0) residual:
 *           r2 = mean_z(r - A z), less its mean (constants are not in the
 *           range of the 2-D operator), is solved approximately on the x-y grid with
 *           the z-averaged operator A2 (a 9-point stencil whose coefficients
 *           are the z-means of the rows' entries grouped by x-y offset) by
 *           ORC_FFT2_ITS steps of flexible GMRES (classical Gram-Schmidt twice,
 *           Givens least squares) preconditioned by the kz = 0 symbol of the
 *           fft preconditioner; the correction, constant along z, is added to z
 *   fft2z   fft2, then zline applied to its residual: z1 = fft2(r),
 *           z = z1 + zline(r - A z1) (the two corrections take complementary
 *           parts of the error: the x-y coupling of the kz = 0 modes and the
 *           density variation along z) */
void orc_pc_apply(int pc, int lhs_type, const int64_t N[3], const double d[3], const double K[3],
                  const double *n, const double *r, double *z);
void orc_pc_apply_deg(int pc, int lhs_type, const int64_t N[3], const double d[3], const double K[3],
                      const double *n, const double *r, double *z, int degree);
/* Solve A phi = b (b consistent, mean-free) starting from phi; right-
 * preconditioned restarted GMRES(m) with two-pass classical Gram-Schmidt,
 * or BiCGStab.  Converged when ||b - A phi|| <= max(rtol ||b||, atol).
 * info[0] = iterations, info[1] = converged (1) or not (0);
 * resid[0] = final true relative residual. */
void orc_solve(int ksp, int pc, int lhs_type, const int64_t N[3], const double d[3],
               const double K[3], const double *n, const double *b, double *phi,
               double rtol, double atol, int64_t max_it, int64_t restart,
               int64_t info[2], double resid[1]);
/* Complex 3-D DFT helper (in place, interleaved re/im, sign -1 forward). */
void orc_fft3d(const int64_t N[3], double *data, int inverse);
/* Host threads of the independent per-ion, per-node and per-line loops; every
 * sum keeps its serial order, so no result depends on the thread count. */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
