/*
 * chy_options.c -- options database and derived parameters.
 *
 * Replaces ProcessOptions / EchoOptions (reference parameters.c:28-654) and the
 * derived grid quantities of SetUpVlasovDM (setup.c:9-49).  The option syntax
 * is PETSc's: "-name value" pairs, a bare "-name" for booleans, names matched
 * case-insensitively, later occurrences overriding earlier ones, plus
 * -options_file <path> and the PETSC_OPTIONS environment variable.
 */
#include "chybrid.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#define CHY_Q  1.6022e-19  /* reference constants.h:5-8 */
#define CHY_MP 1.6726e-27
#define CHY_ME 9.1094e-31
#define CHY_KB 1.3807e-23

/* ---- a small options database ------------------------------------------- */
typedef struct {
  char **key;
  char **val; /* NULL for a bare flag */
  int n, cap;
} optdb;

static void db_add(optdb *db, const char *k, const char *v)
{
  if (db->n == db->cap) {
    db->cap = db->cap ? 2 * db->cap : 64;
    db->key = realloc(db->key, sizeof(char *) * db->cap);
    db->val = realloc(db->val, sizeof(char *) * db->cap);
  }
  db->key[db->n] = strdup(k);
  db->val[db->n] = v ? strdup(v) : NULL;
  db->n++;
}

static void db_free(optdb *db)
{
  for (int i = 0; i < db->n; i++) { free(db->key[i]); free(db->val[i]); }
  free(db->key); free(db->val);
  memset(db, 0, sizeof(*db));
}

/* A token is an option name when it starts with '-' and is not a number. */
static int is_name(const char *s)
{
  if (!s || s[0] != '-') return 0;
  if (isdigit((unsigned char)s[1]) || s[1] == '.') return 0;
  return s[1] != '\0';
}

static void db_parse_tokens(optdb *db, int n, char **tok);

static void db_parse_file(optdb *db, const char *path)
{
  FILE *f = fopen(path, "r");
  if (!f) return;
  char line[8192];
  char *toks[512];
  while (fgets(line, sizeof line, f)) {
    char *h = strchr(line, '#');
    if (h) *h = '\0';
    int n = 0;
    for (char *t = strtok(line, " \t\r\n"); t && n < 512; t = strtok(NULL, " \t\r\n")) toks[n++] = t;
    db_parse_tokens(db, n, toks);
  }
  fclose(f);
}

static void db_parse_tokens(optdb *db, int n, char **tok)
{
  for (int i = 0; i < n; i++) {
    if (!is_name(tok[i])) continue;
    const char *v = (i + 1 < n && !is_name(tok[i + 1])) ? tok[i + 1] : NULL;
    if (strcasecmp(tok[i], "-options_file") == 0 && v) db_parse_file(db, v);
    else db_add(db, tok[i], v);
    if (v) i++;
  }
}

/* Last occurrence wins. */
static const char *db_get(const optdb *db, const char *name, int *found)
{
  for (int i = db->n - 1; i >= 0; i--) {
    if (strcasecmp(db->key[i], name) == 0) {
      *found = 1;
      return db->val[i];
    }
  }
  *found = 0;
  return NULL;
}

typedef struct { char *buf; size_t len; int status; } errbuf;

static void emsg(errbuf *e, int status, const char *fmt, ...)
{
  if (status && !e->status) e->status = status;
  if (!e->buf || !e->len) return;
  size_t used = strlen(e->buf);
  if (used + 2 >= e->len) return;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(e->buf + used, e->len - used, fmt, ap);
  va_end(ap);
}

static int get_real(const optdb *db, const char *name, double *out, errbuf *e)
{
  int f;
  const char *v = db_get(db, name, &f);
  if (!f) return 0;
  char *end = NULL;
  const double x = v ? strtod(v, &end) : NAN;
  if (!v || end == v || *end) { emsg(e, CHY_ERR_ARG_WRONG, "bad real value for %s\n", name); return 0; }
  *out = x;
  return 1;
}

static int get_int(const optdb *db, const char *name, int64_t *out, errbuf *e)
{
  int f;
  const char *v = db_get(db, name, &f);
  if (!f) return 0;
  char *end = NULL;
  const long long x = v ? strtoll(v, &end, 10) : 0;
  if (!v || end == v || *end) { emsg(e, CHY_ERR_ARG_WRONG, "bad integer value for %s\n", name); return 0; }
  *out = (int64_t)x;
  return 1;
}

static int get_bool(const optdb *db, const char *name, int32_t *out, errbuf *e)
{
  int f;
  const char *v = db_get(db, name, &f);
  if (!f) return 0;
  if (!v || !strcasecmp(v, "1") || !strcasecmp(v, "true") || !strcasecmp(v, "yes") || !strcasecmp(v, "on")) *out = 1;
  else if (!strcasecmp(v, "0") || !strcasecmp(v, "false") || !strcasecmp(v, "no") || !strcasecmp(v, "off")) *out = 0;
  else { emsg(e, CHY_ERR_ARG_WRONG, "bad boolean value for %s\n", name); return 0; }
  return 1;
}

static int get_enum(const optdb *db, const char *name, const char *const *names, int32_t *out, errbuf *e)
{
  int f;
  const char *v = db_get(db, name, &f);
  if (!f) return 0;
  for (int i = 0; v && names[i]; i++) {
    if (!strcasecmp(v, names[i])) { *out = i; return 1; }
  }
  emsg(e, CHY_ERR_ARG_WRONG, "unknown value '%s' for %s\n", v ? v : "", name);
  return 0;
}

/* Reference enum name tables, parameters.c:8-25. */
static const char *const RHSNames[] = {"constant", "sinusoidal", "full", NULL};
static const char *const LHSNames[] = {"identity", "laplacian", "full", NULL};
static const char *const DensityNames[] = {"flat-sobol", "flat-reverse", "flat-normal", "uniform",
                                           "sinusoidal", "gaussian", "lattice", NULL};
static const char *const BCNames[] = {"periodic", "injection", "reflection", "dirichlet", "neumann", NULL};
static const char *const PertNames[] = {"none", "sinusoidal-x", "uniform", "gaussian", NULL};
static const char *const PrecNames[] = {"fp64", "mixed", NULL};
/* PETSc KSPGMRESCGSRefinementType names */
static const char *const CgsRefineNames[] = {"REFINE_NEVER", "REFINE_IFNEEDED", "REFINE_ALWAYS", NULL};
static const char *const GuessNames[] = {"previous", "zero", "extrapolate", NULL};

int chy_params_default(chy_params *p)
{
  memset(p, 0, sizeof(*p));
  for (int a = 0; a < 3; a++) {
    p->N[a] = -1;
    p->p0[a] = 0.0;
    p->p1[a] = 1.0;
    p->bc[a][0] = p->bc[a][1] = CHY_BC_PERIODIC;
    p->flux_scale[a] = 1.0;
  }
  p->qe = -CHY_Q;
  p->me = CHY_ME;
  p->nue = 1.0;
  p->qi = CHY_Q;
  p->mi = CHY_MP;
  p->nui = 1.0;
  p->mn = -1.0;      /* = mi unless given */
  p->Tn = -1.0;      /* = Ti unless given */
  p->Np = -1;
  p->Nt = 1;
  p->dt = -1.0;      /* = 1/nui unless given */
  p->lhs_type = CHY_LHS_FULL;
  p->rhs_type = CHY_RHS_FULL;
  p->density_type = CHY_DENSITY_FLAT_SOBOL;
  p->ksp_type = CHY_KSP_GMRES;
  p->pc_type = CHY_PC_FFT2Z;
  p->ksp_rtol = 1e-5;      /* PETSc KSP defaults */
  p->ksp_atol = 1e-50;
  p->ksp_max_it = 10000;
  p->ksp_restart = 30;
  p->pc_cheby_degree = 4;
  p->ksp_inner_rtol = 2e-5;
  p->collision_refusal_limit = 10;  /* particles.c:603 */
  p->ksp_gmres_cgs_refinement = CHY_CGS_REFINE_NEVER;  /* PETSc default */
  p->precision = CHY_PREC_FP64;
  p->output_interval = 1;
  p->perturbation = CHY_PERT_NONE;
  snprintf(p->output, sizeof p->output, "results");
  return CHY_OK;
}

static void get_vec3(const optdb *db, const char *base, double v[3], errbuf *e)
{
  /* common value for all components, then per-component overrides
   * (parameters.c:221-245 pattern) */
  char name[64];
  double x;
  if (get_real(db, base, &x, e)) v[0] = v[1] = v[2] = x;
  const char *ax = "xyz";
  for (int a = 0; a < 3; a++) {
    snprintf(name, sizeof name, "%s%c", base, ax[a]);
    get_real(db, name, &v[a], e);
  }
}

static double norm3(const double v[3]) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

int chy_params_derive(chy_params *p, char *err, size_t errlen)
{
  errbuf e = {err, errlen, 0};
  for (int a = 0; a < 3; a++) {
    if (p->N[a] <= 0) p->N[a] = 7; /* DMDACreate3d default used by setup.c:9-11 */
    p->L[a] = p->p1[a] - p->p0[a];
    p->d[a] = p->L[a] / (double)p->N[a];
    if (!(p->L[a] > 0.0)) emsg(&e, CHY_ERR_ARG_WRONG, "non-positive domain length on axis %d\n", a);
    if (p->N[a] < 3) emsg(&e, CHY_ERR_ARG_RANGE, "the periodic 19-point operator needs >= 3 nodes per axis\n");
    /* parameters.c:481-506 */
    const int lo = p->bc[a][0] == CHY_BC_PERIODIC, hi = p->bc[a][1] == CHY_BC_PERIODIC;
    if (lo != hi) emsg(&e, CHY_ERR_ARG_INCOMP, "Inconsistent %c-axis boundary conditions: {%s, %s}\n",
                       "xyz"[a], BCNames[p->bc[a][0]], BCNames[p->bc[a][1]]);
    else if (!lo) emsg(&e, CHY_ERR_ARG_WRONG, "Unsupported boundary conditions on the %c axis\n", "xyz"[a]);
  }
  const int64_t ncell = p->N[0] * p->N[1] * p->N[2];
  if (p->ppc > 0) {
    if (p->Np > 0 && p->Np != p->ppc * ncell)
      emsg(&e, CHY_ERR_ARG_INCOMP, "-Np %lld differs from -ppc x cells = %lld\n",
           (long long)p->Np, (long long)(p->ppc * ncell));
    p->Np = p->ppc * ncell;
  }
  if (p->Np < 0) p->Np = ncell; /* setup.c:43-45 */
  if (p->Np > 0xFFFFFFFFll) emsg(&e, CHY_ERR_ARG_RANGE, "at most 2^32-1 particles are supported\n");
  if (p->mn < 0.0) p->mn = p->mi;
  if (p->dt <= 0.0) p->dt = 1.0 / p->nui; /* parameters.c:412-418 */
  /* gyrofrequencies and magnetisation, parameters.c:507-520 */
  for (int a = 0; a < 3; a++) {
    p->Omega_e[a] = fabs(p->qe * p->B0[a] / p->me);
    p->Omega_i[a] = fabs(p->qi * p->B0[a] / p->mi);
    p->kappa_e[a] = p->Omega_e[a] / p->nue;
    p->kappa_i[a] = p->Omega_i[a] / p->nui;
  }
  p->vn0_r = norm3(p->vn0);
  /* temperatures and thermal speeds, parameters.c:525-546 (vT per component
   * is sqrt(2 kB T / m) / 3, kept as the flag's meaning) */
  if (p->veT[0] != 0.0 || p->veT[1] != 0.0 || p->veT[2] != 0.0) {
    p->veT_r = norm3(p->veT);
    p->Te = (0.5 * p->me / CHY_KB) * (p->veT_r * p->veT_r);
  } else {
    const double t = sqrt(2.0 * CHY_KB * p->Te / p->me) / 3.0;
    p->veT[0] = p->veT[1] = p->veT[2] = t;
    p->veT_r = norm3(p->veT);
  }
  if (p->viT[0] != 0.0 || p->viT[1] != 0.0 || p->viT[2] != 0.0) {
    p->viT_r = norm3(p->viT);
    p->Ti = (0.5 * p->mi / CHY_KB) * (p->viT_r * p->viT_r);
  } else {
    const double t = sqrt(2.0 * CHY_KB * p->Ti / p->mi) / 3.0;
    p->viT[0] = p->viT[1] = p->viT[2] = t;
    p->viT_r = norm3(p->viT);
  }
  if (p->Tn == -1.0) p->Tn = p->Ti; /* parameters.c:548-551 */
  p->vnT = (p->mn > 0.0) ? sqrt(2.0 * CHY_KB * p->Tn / p->mn) : 0.0;
  p->gamma_e = 1.0; /* parameters.c:556 */
  if (p->precision != CHY_PREC_FP64 && p->precision != CHY_PREC_MIXED)
    emsg(&e, CHY_ERR_ARG_WRONG, "unknown precision\n");
  if (p->ksp_restart < 1) p->ksp_restart = 30;
  if (p->pc_cheby_degree < 1) p->pc_cheby_degree = 4;
  if (!(p->ksp_inner_rtol > 0.0 && p->ksp_inner_rtol < 1.0)) p->ksp_inner_rtol = 2e-5;
  return e.status;
}

int chy_params_from_options(int argc, const char *const *argv, chy_params *p, char *err, size_t errlen)
{
  optdb db = {0};
  errbuf e = {err, errlen, 0};
  if (err && errlen) err[0] = '\0';
  chy_params_default(p);
  const char *env = getenv("PETSC_OPTIONS");
  if (env) {
    char *copy = strdup(env);
    char *toks[1024];
    int n = 0;
    for (char *t = strtok(copy, " \t\n"); t && n < 1024; t = strtok(NULL, " \t\n")) toks[n++] = t;
    db_parse_tokens(&db, n, toks);
    free(copy);
  }
  db_parse_tokens(&db, argc, (char **)argv);

  int f;
  const char *s;
  if ((s = db_get(&db, "--input", &f)) && f) snprintf(p->input, sizeof p->input, "%s", s);
  if ((s = db_get(&db, "--output", &f)) && f) snprintf(p->output, sizeof p->output, "%s", s);
  get_bool(&db, "--view-lhs", &p->view_lhs, &e);
  get_enum(&db, "--rhs-type", RHSNames, &p->rhs_type, &e);
  get_enum(&db, "--lhs-type", LHSNames, &p->lhs_type, &e);
  const int density_given = get_enum(&db, "--density-type", DensityNames, &p->density_type, &e);
  const char *nn[3] = {"-Nx", "-Ny", "-Nz"};
  const char *n0[3] = {"-x0", "-y0", "-z0"}, *n1[3] = {"-x1", "-y1", "-z1"};
  const char *b0[3] = {"--x0-bc", "--y0-bc", "--z0-bc"}, *b1[3] = {"--x1-bc", "--y1-bc", "--z1-bc"};
  for (int a = 0; a < 3; a++) {
    int64_t v;
    if (get_int(&db, nn[a], &v, &e)) {
      if (v < 0) emsg(&e, 0, "Warning: Ignoring negative value for %s: %lld\n", nn[a] + 1, (long long)v);
      else p->N[a] = v;
    }
    get_real(&db, n0[a], &p->p0[a], &e);
    get_real(&db, n1[a], &p->p1[a], &e);
    get_enum(&db, b0[a], BCNames, &p->bc[a][0], &e);
    get_enum(&db, b1[a], BCNames, &p->bc[a][1], &e);
  }
  get_int(&db, "-Np", &p->Np, &e);
  get_real(&db, "-qi", &p->qi, &e);
  get_real(&db, "-mi", &p->mi, &e);
  get_real(&db, "-nue", &p->nue, &e);
  get_real(&db, "-nui", &p->nui, &e);
  get_vec3(&db, "-ve0", p->ve0, &e);
  get_vec3(&db, "-vi0", p->vi0, &e);
  get_vec3(&db, "-veT", p->veT, &e);
  get_real(&db, "-Te", &p->Te, &e);
  get_vec3(&db, "-viT", p->viT, &e);
  get_real(&db, "-Ti", &p->Ti, &e);
  get_real(&db, "-mn", &p->mn, &e);
  get_real(&db, "-Tn", &p->Tn, &e);
  get_vec3(&db, "-vn0", p->vn0, &e);
  get_real(&db, "-B0x", &p->B0[0], &e);
  get_real(&db, "-B0y", &p->B0[1], &e);
  get_real(&db, "-B0z", &p->B0[2], &e);
  get_real(&db, "-E0x", &p->E0[0], &e);
  get_real(&db, "-E0y", &p->E0[1], &e);
  get_real(&db, "-E0z", &p->E0[2], &e);
  get_int(&db, "-Nt", &p->Nt, &e);
  get_real(&db, "-dt", &p->dt, &e);
  double fs;
  if (get_real(&db, "--flux-scale", &fs, &e)) p->flux_scale[0] = p->flux_scale[1] = p->flux_scale[2] = fs;
  get_real(&db, "--x-flux-scale", &p->flux_scale[0], &e);
  get_real(&db, "--y-flux-scale", &p->flux_scale[1], &e);
  get_real(&db, "--z-flux-scale", &p->flux_scale[2], &e);

  /* build additions */
  int64_t iv;
  if (get_int(&db, "-seed", &iv, &e)) p->seed = (uint64_t)iv;
  get_int(&db, "-ppc", &p->ppc, &e);
  get_int(&db, "-diagnostics_interval", &p->diagnostics_interval, &e);
  /* -ppc selects the per-cell lattice placement (a build addition) */
  if (p->ppc > 0) {
    if (density_given && p->density_type != CHY_DENSITY_LATTICE)
      emsg(&e, CHY_ERR_ARG_INCOMP, "-ppc places ions per cell; it does not combine with --density-type %s\n",
           DensityNames[p->density_type]);
    p->density_type = CHY_DENSITY_LATTICE;
  }
  get_enum(&db, "--perturbation", PertNames, &p->perturbation, &e);
  get_real(&db, "--perturbation-amplitude", &p->perturbation_amplitude, &e);
  get_enum(&db, "-precision", PrecNames, &p->precision, &e);
  get_int(&db, "-output_interval", &p->output_interval, &e);

  /* PETSc KSP / PC subset */
  if ((s = db_get(&db, "-ksp_type", &f)) && f) {
    if (!strcasecmp(s, "gmres") || !strcasecmp(s, "fgmres") || !strcasecmp(s, "lgmres")) p->ksp_type = CHY_KSP_GMRES;
    else if (!strcasecmp(s, "bcgs") || !strcasecmp(s, "bcgsl") || !strcasecmp(s, "ibcgs")) p->ksp_type = CHY_KSP_BCGS;
    else if (!strcasecmp(s, "preonly") || !strcasecmp(s, "richardson") || !strcasecmp(s, "cg"))
      emsg(&e, CHY_ERR_SUP, "-ksp_type %s is not supported on the GPU solver\n", s);
    else emsg(&e, 0, "Warning: -ksp_type %s mapped to gmres\n", s);
  }
  if ((s = db_get(&db, "-pc_type", &f)) && f) {
    if (!strcasecmp(s, "none")) p->pc_type = CHY_PC_NONE;
    else if (!strcasecmp(s, "jacobi") || !strcasecmp(s, "pbjacobi")) p->pc_type = CHY_PC_JACOBI;
    else if (!strcasecmp(s, "fft")) p->pc_type = CHY_PC_FFT;
    else if (!strcasecmp(s, "fft2")) p->pc_type = CHY_PC_FFT2;
    else if (!strcasecmp(s, "fft2z")) p->pc_type = CHY_PC_FFT2Z;
    else if (!strcasecmp(s, "zline") || !strcasecmp(s, "bjacobi") || !strcasecmp(s, "line"))
      p->pc_type = CHY_PC_ZLINE;
    else if (!strcasecmp(s, "cheby") || !strcasecmp(s, "chebyshev")) p->pc_type = CHY_PC_CHEBY;
    else if (!strcasecmp(s, "mg") || !strcasecmp(s, "gamg") || !strcasecmp(s, "hypre") || !strcasecmp(s, "ml"))
      p->pc_type = CHY_PC_MG;  /* PETSc's multigrid preconditioners: x-y semicoarsening multigrid */
    else emsg(&e, 0, "Warning: -pc_type %s mapped to the default GPU preconditioner (fft2z)\n", s);
  }
  get_real(&db, "-ksp_rtol", &p->ksp_rtol, &e);
  get_real(&db, "-ksp_atol", &p->ksp_atol, &e);
  get_int(&db, "-ksp_max_it", &p->ksp_max_it, &e);
  get_int(&db, "-ksp_gmres_restart", &p->ksp_restart, &e);
  get_real(&db, "-ksp_inner_rtol", &p->ksp_inner_rtol, &e);
  get_enum(&db, "-ksp_gmres_cgs_refinement_type", CgsRefineNames, &p->ksp_gmres_cgs_refinement, &e);
  get_enum(&db, "-ksp_initial_guess", GuessNames, &p->ksp_initial_guess, &e);
  {
    /* the reference aborts a step at its 10th refused scattering; with ~1e9
       ions and tens of millions of scatterings per step a few refusals are
       routine, so the limit is a flag (0: never abort) */
    int64_t lim = p->collision_refusal_limit;
    if (get_int(&db, "--collision-refusal-limit", &lim, &e)) {
      if (lim < 0) emsg(&e, CHY_ERR_ARG_RANGE, "--collision-refusal-limit must be >= 0\n");
      else p->collision_refusal_limit = lim;
    }
  }
  {
    int64_t deg = p->pc_cheby_degree;
    if (get_int(&db, "-pc_cheby_degree", &deg, &e)) {
      if (deg < 1 || deg > 64) emsg(&e, CHY_ERR_ARG_RANGE, "-pc_cheby_degree must be in [1, 64]\n");
      else p->pc_cheby_degree = (int32_t)deg;
    }
  }
  int32_t bv;
  if (get_bool(&db, "-ksp_monitor", &bv, &e)) p->ksp_monitor = bv;
  if (get_bool(&db, "-ksp_monitor_short", &bv, &e)) p->ksp_monitor = bv;
  get_bool(&db, "-ksp_converged_reason", &p->ksp_converged_reason, &e);
  get_bool(&db, "-ksp_error_if_not_converged", &p->ksp_error_if_not_converged, &e);

  db_free(&db);
  const int st = chy_params_derive(p, err ? err + strlen(err) : NULL, err ? errlen - strlen(err) : 0);
  return e.status ? e.status : st;
}

int chy_write_options_txt(const chy_params *c, int argc, const char *const *argv, const char *path)
{
  FILE *f = fopen(path, "w");
  if (!f) return CHY_ERR_LIB;
  /* PetscOptionsView layout: one "-key value" line per option */
  fprintf(f, "#PETSc Option Table entries:\n");
  for (int i = 0; i < argc; i++) {
    if (!is_name(argv[i])) continue;
    if (i + 1 < argc && !is_name(argv[i + 1])) { fprintf(f, "%s %s\n", argv[i], argv[i + 1]); i++; }
    else fprintf(f, "%s\n", argv[i]);
  }
  fprintf(f, "#End of PETSc Option Table entries\n");
  /* parameters.c:580-647 */
  fprintf(f, "\nParameter Values\n");
  fprintf(f, "----------------\n");
  fprintf(f, "Nt = %lld\n", (long long)c->Nt);
  fprintf(f, "dt = %f [s]\n", c->dt);
  fprintf(f, "Nx = %lld\n", (long long)c->N[0]);
  fprintf(f, "Ny = %lld\n", (long long)c->N[1]);
  fprintf(f, "Nz = %lld\n", (long long)c->N[2]);
  fprintf(f, "x0 = %f [m]\n", c->p0[0]);
  fprintf(f, "y0 = %f [m]\n", c->p0[1]);
  fprintf(f, "z0 = %f [m]\n", c->p0[2]);
  fprintf(f, "x1 = %f [m]\n", c->p1[0]);
  fprintf(f, "y1 = %f [m]\n", c->p1[1]);
  fprintf(f, "z1 = %f [m]\n", c->p1[2]);
  fprintf(f, "Lx = %f [m]\n", c->L[0]);
  fprintf(f, "Ly = %f [m]\n", c->L[1]);
  fprintf(f, "Lz = %f [m]\n", c->L[2]);
  fprintf(f, "dx = %f [m]\n", c->d[0]);
  fprintf(f, "dy = %f [m]\n", c->d[1]);
  fprintf(f, "dz = %f [m]\n", c->d[2]);
  fprintf(f, "xBC = {%s, %s}\n", BCNames[c->bc[0][0]], BCNames[c->bc[0][1]]);
  fprintf(f, "yBC = {%s, %s}\n", BCNames[c->bc[1][0]], BCNames[c->bc[1][1]]);
  fprintf(f, "zBC = {%s, %s}\n", BCNames[c->bc[2][0]], BCNames[c->bc[2][1]]);
  fprintf(f, "Np = %lld\n", (long long)c->Np);
  fprintf(f, "B0x = %g [T]\n", c->B0[0]);
  fprintf(f, "B0y = %g [T]\n", c->B0[1]);
  fprintf(f, "B0z = %g [T]\n", c->B0[2]);
  fprintf(f, "E0x = %g [N/C]\n", c->E0[0]);
  fprintf(f, "E0y = %g [N/C]\n", c->E0[1]);
  fprintf(f, "E0z = %g [N/C]\n", c->E0[2]);
  fprintf(f, "qe = %e [C]\n", c->qe);
  fprintf(f, "me = %e [kg]\n", c->me);
  fprintf(f, "nue = %g [s^-1]\n", c->nue);
  fprintf(f, "ve0x = %f [m/s]\n", c->ve0[0]);
  fprintf(f, "ve0y = %f [m/s]\n", c->ve0[1]);
  fprintf(f, "ve0z = %f [m/s]\n", c->ve0[2]);
  fprintf(f, "veTx = %f [m/s]\n", c->veT[0]);
  fprintf(f, "veTy = %f [m/s]\n", c->veT[1]);
  fprintf(f, "veTz = %f [m/s]\n", c->veT[2]);
  fprintf(f, "Te = %f [K]\n", c->Te);
  fprintf(f, "gamma_e = %f\n", c->gamma_e);
  fprintf(f, "Omega_ex = %g [s^-1]\n", c->Omega_e[0]);
  fprintf(f, "Omega_ey = %g [s^-1]\n", c->Omega_e[1]);
  fprintf(f, "Omega_ez = %g [s^-1]\n", c->Omega_e[2]);
  fprintf(f, "kappa_ex = %g\n", c->kappa_e[0]);
  fprintf(f, "kappa_ey = %g\n", c->kappa_e[1]);
  fprintf(f, "kappa_ez = %g\n", c->kappa_e[2]);
  fprintf(f, "qi = %e [C]\n", c->qi);
  fprintf(f, "mi = %e [kg]\n", c->mi);
  fprintf(f, "nui = %g [s^-1]\n", c->nui);
  fprintf(f, "vi0x = %f [m/s]\n", c->vi0[0]);
  fprintf(f, "vi0y = %f [m/s]\n", c->vi0[1]);
  fprintf(f, "vi0z = %f [m/s]\n", c->vi0[2]);
  fprintf(f, "viTx = %f [m/s]\n", c->viT[0]);
  fprintf(f, "viTy = %f [m/s]\n", c->viT[1]);
  fprintf(f, "viTz = %f [m/s]\n", c->viT[2]);
  fprintf(f, "Ti = %f [K]\n", c->Ti);
  fprintf(f, "Omega_ix = %g [s^-1]\n", c->Omega_i[0]);
  fprintf(f, "Omega_iy = %g [s^-1]\n", c->Omega_i[1]);
  fprintf(f, "Omega_iz = %g [s^-1]\n", c->Omega_i[2]);
  fprintf(f, "kappa_ix = %g\n", c->kappa_i[0]);
  fprintf(f, "kappa_iy = %g\n", c->kappa_i[1]);
  fprintf(f, "kappa_iz = %g\n", c->kappa_i[2]);
  fprintf(f, "mn = %e\n", c->mn);
  fprintf(f, "Tn = %f\n", c->Tn);
  fprintf(f, "vnT = %f\n", c->vnT);
  fprintf(f, "vn0x = %f\n", c->vn0[0]);
  fprintf(f, "vn0y = %f\n", c->vn0[1]);
  fprintf(f, "vn0z = %f\n", c->vn0[2]);
  fclose(f);
  return CHY_OK;
}
