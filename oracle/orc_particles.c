/*
 * orc_particles.c -- oracle (test infrastructure only): counter-based RNG,
 * Boris push, ion-neutral collisions, periodic move, stable cell sort and
 * seeded initial state.  FP64 restatement; see orc.h for the role of this file.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ------------------- */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r > 0) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
    const uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    const uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double orc_u24(uint32_t w) { return ((double)(w >> 8) + 0.5) * (1.0 / 16777216.0); }

static void philox_u(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint64_t seed, double u[4])
{
  const uint32_t ctr[4] = {a, b, c, d};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  orc_philox4x32_10(ctr, key, o);
  for (int q = 0; q < 4; q++) u[q] = orc_u24(o[q]);
}

/* RNG streams: ctr = {id, step, stream, block}. */
enum { STREAM_INIT = 0, STREAM_COLLIDE_A = 1, STREAM_COLLIDE_B = 2, STREAM_WEIGHT = 3 };

double orc_wrap(double x, double p0, double p1, double L)
{
  /* far outside (more than one period): reduce with floor first */
  if (!(x >= p0 - L && x < p1 + L)) {
    if (!isfinite(x)) return x;
    x = x - floor((x - p0) / L) * L;
  }
  if (x < p0) {
    x += L;
    if (x >= p1) x = p0;
  } else if (x >= p1) {
    x -= L;
    if (x < p0) x = p0;
  }
  return x;
}

double orc_mean_rel_speed(const double v[3], const double vn0[3], double vnT)
{
  const double ux = v[0] - vn0[0], uy = v[1] - vn0[1], uz = v[2] - vn0[2];
  const double mu = sqrt(ux * ux + uy * uy + uz * uz);
  if (!(vnT > 0.0)) return mu;
  const double lam = mu / vnT;
  const double c = 0.79788456080286535588; /* sqrt(2/pi) */
  if (lam < 1e-3) return vnT * c * (2.0 + lam * lam / 3.0);
  return vnT * (c * exp(-0.5 * lam * lam) + (lam + 1.0 / lam) * erf(lam * 0.70710678118654752440));
}

/* Collision normaliser of a state: sum over the ions of the mean relative
 * speed (the S that the next step's collision probabilities divide by). */
double orc_normaliser(int64_t np, const double *vx, const double *vy, const double *vz, const double vn0[3],
                      double vnT)
{
  double S = 0.0;
  for (int64_t p = 0; p < np; p++) {
    const double v[3] = {vx[p], vy[p], vz[p]};
    S += orc_mean_rel_speed(v, vn0, vnT);
  }
  return S;
}

int64_t orc_collision_count(int64_t Np, double nui, double dt, double viT,
                            const double vi0[3], const double vn0[3], double vnT,
                            double mi, double mn)
{
  const double fc = nui * dt;
  const double dvx = vi0[0] - vn0[0], dvy = vi0[1] - vn0[1], dvz = vi0[2] - vn0[2];
  const double vrm = 4.0 * viT + sqrt(dvx * dvx + dvy * dvy + dvz * dvz);
  const double vn0r = sqrt(vn0[0] * vn0[0] + vn0[1] * vn0[1] + vn0[2] * vn0[2]);
  const double den = (vnT + vn0r) * sqrt(mn / mi);
  if (!(den > 0.0)) return Np;
  const double x = ((double)Np * fc * vrm) / den;
  if (!(x < (double)Np)) return Np;
  if (!(x > 0.0)) return 0;
  return (int64_t)x;
}

/* Scattering of one ion on a Maxwellian neutral, particles.c:537-611 with a
 * per-ion Bernoulli trial (probability Nc |v_r| / S) in place of the serial
 * "pick until Nc accepted" loop, and keyed random numbers.  Returns 1 when
 * accepted, 2 when refused (final speed > 10 vT), 0 otherwise; *capped set
 * when the probability had to be clipped at 1. */
static int collide_one(const orc_push_params *P, uint32_t pid, double S, double v[3], int *capped)
{
  /* One Philox block gives the neutral draw (Box-Muller on the top 24 bits
   * of the four words) and the decision uniform (from their low bytes). */
  const uint32_t ctr[4] = {pid, P->step, STREAM_COLLIDE_A, 0};
  const uint32_t key[2] = {(uint32_t)P->seed, (uint32_t)(P->seed >> 32)};
  uint32_t a[4];
  orc_philox4x32_10(ctr, key, a);
  double ua[4], ub[4];
  for (int q = 0; q < 4; q++) ua[q] = orc_u24(a[q]);
  const double udec = orc_u24(((a[0] & 0xffu) | ((a[1] & 0xffu) << 8) | ((a[2] & 0xffu) << 16)) << 8);
  const double r1 = sqrt(-2.0 * log(ua[0])), a1 = 2.0 * M_PI * ua[1];
  const double r2 = sqrt(-2.0 * log(ua[2])), a2 = 2.0 * M_PI * ua[3];
  const double vnx = P->vn0[0] + P->vnT * (r1 * cos(a1));
  const double vny = P->vn0[1] + P->vnT * (r1 * sin(a1));
  const double vnz = P->vn0[2] + P->vnT * (r2 * cos(a2));
  const double vrx = v[0] - vnx, vry = v[1] - vny, vrz = v[2] - vnz;
  const double vrr = sqrt(vrx * vrx + vry * vry + vrz * vrz);
  double p = (S > 0.0) ? ((double)P->Nc * vrr) / S : 0.0;
  *capped = 0;
  if (p > 1.0) { p = 1.0; *capped = 1; }
  if (!(udec < p)) return 0;
  philox_u(pid, P->step, STREAM_COLLIDE_B, 0, P->seed, ub);
  const double mi = P->mi, mn = P->mn, M = mn + mi;
  const double vcx = (v[0] * mi + vnx * mn) / M;
  const double vcy = (v[1] * mi + vny * mn) / M;
  const double vcz = (v[2] * mi + vnz * mn) / M;
  const double ex = v[0] - vcx, ey = v[1] - vcy, ez = v[2] - vcz;
  const double vcr = sqrt(ex * ex + ey * ey + ez * ez);
  const double uz = 2.0 * ub[0] - 1.0;
  const double uperp = sqrt(1.0 - uz * uz);
  const double uphi = 2.0 * M_PI * ub[1];
  const double ux = uperp * cos(uphi), uy = uperp * sin(uphi);
  const double costht = vrz / vrr;
  const double sintht = sqrt(1.0 - costht * costht);
  double cosphi = 1.0, sinphi = 0.0;
  if (sintht > 1e-12) { /* intended guard for an axial v_r (particles.c:570-571) */
    cosphi = vrx / (vrr * sintht);
    sinphi = vry / (vrr * sintht);
  }
  const double ux1 = uz * sintht + ux * costht, uy1 = uy, uz1 = uz * costht - ux * sintht;
  const double ux2 = ux1 * cosphi - uy1 * sinphi, uy2 = ux1 * sinphi + uy1 * cosphi, uz2 = uz1;
  const double vfx = vcx + vcr * ux2, vfy = vcy + vcr * uy2, vfz = vcz + vcr * uz2;
  const double vfr = sqrt(vfx * vfx + vfy * vfy + vfz * vfz);
  if (vfr / P->viT > 10.0) return 2;
  v[0] = vfx; v[1] = vfy; v[2] = vfz;
  return 1;
}

void orc_push(const orc_push_params *P, const double *ex, const double *ey, const double *ez,
              int64_t np, double *x, double *y, double *z,
              double *vx, double *vy, double *vz, const uint32_t *pid,
              double S, int32_t *key, double *S_next, int64_t stats[3])
{
  /* Boris constants, particles.c:374-384 (t = q B dt / 2m, s = 2t/(1+t.t)). */
  const double tau = 0.5 * P->qm * P->dt;
  const double t0 = tau * P->B0[0], t1 = tau * P->B0[1], t2 = tau * P->B0[2];
  const double tt = t0 * t0 + t1 * t1 + t2 * t2;
  const double s0 = 2.0 * t0 / (1 + tt), s1 = 2.0 * t1 / (1 + tt), s2 = 2.0 * t2 / (1 + tt);
  /* independent ions: threads change nothing (the counts are integers and
   * the normaliser is summed in ion order after the loop) */
  double *mrs = (double *)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
  int64_t acc = 0, ref = 0, cap = 0;
#pragma omp parallel for schedule(static) reduction(+ : acc, ref, cap) if (np > 16384)
  for (int64_t p = 0; p < np; p++) {
    double E[3];
    orc_gather(P->N, P->p0, P->d, ex, ey, ez, x[p], y[p], z[p], E);
    /* Boris rotation with the intended per-component v-minus and cross
     * products (particles.c:428-447, vectors.c:153-162). */
    const double vm0 = vx[p] + tau * E[0], vm1 = vy[p] + tau * E[1], vm2 = vz[p] + tau * E[2];
    const double vp0 = vm0 + (vm1 * t2 - vm2 * t1);
    const double vp1 = vm1 + (vm2 * t0 - vm0 * t2);
    const double vp2 = vm2 + (vm0 * t1 - vm1 * t0);
    const double vq0 = vm0 + (vp1 * s2 - vp2 * s1);
    const double vq1 = vm1 + (vp2 * s0 - vp0 * s2);
    const double vq2 = vm2 + (vp0 * s1 - vp1 * s0);
    double v[3] = {vq0 + tau * E[0], vq1 + tau * E[1], vq2 + tau * E[2]};
    if (P->collide) {
      int capped = 0;
      const int r = collide_one(P, pid[p], S, v, &capped);
      if (r == 1) acc++;
      if (r == 2) ref++;
      cap += capped;
    }
    vx[p] = v[0]; vy[p] = v[1]; vz[p] = v[2];
    /* Move with stored result (intended; particles.c:679-704). */
    x[p] = orc_wrap(x[p] + v[0] * P->dt, P->p0[0], P->p1[0], P->L[0]);
    y[p] = orc_wrap(y[p] + v[1] * P->dt, P->p0[1], P->p1[1], P->L[1]);
    z[p] = orc_wrap(z[p] + v[2] * P->dt, P->p0[2], P->p1[2], P->L[2]);
    int64_t i, j, k;
    double f;
    orc_cell_coord(x[p], P->p0[0], P->d[0], P->N[0], &i, &f);
    orc_cell_coord(y[p], P->p0[1], P->d[1], P->N[1], &j, &f);
    orc_cell_coord(z[p], P->p0[2], P->d[2], P->N[2], &k, &f);
    key[p] = (int32_t)(i + P->N[0] * (j + P->N[1] * k));
    mrs[p] = orc_mean_rel_speed(v, P->vn0, P->vnT);
  }
  double Sn = 0.0;
  for (int64_t p = 0; p < np; p++) Sn += mrs[p];
  free(mrs);
  stats[0] = acc; stats[1] = ref; stats[2] = cap;
  *S_next = Sn;
}

void orc_sort_perm(int64_t ncell, int64_t np, const int32_t *key, int64_t *perm, int64_t *offsets)
{
  int64_t *cur = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
  for (int64_t c = 0; c <= ncell; c++) offsets[c] = 0;
  for (int64_t p = 0; p < np; p++) offsets[key[p] + 1]++;
  for (int64_t c = 0; c < ncell; c++) offsets[c + 1] += offsets[c];
  for (int64_t c = 0; c < ncell; c++) cur[c] = offsets[c];
  for (int64_t p = 0; p < np; p++) perm[cur[key[p]]++] = p;
  free(cur);
}

void orc_permute_d(int64_t np, const int64_t *perm, const double *in, double *out)
{
#pragma omp parallel for schedule(static) if (np > 65536)
  for (int64_t s = 0; s < np; s++) out[s] = in[perm[s]];
}

void orc_permute_u(int64_t np, const int64_t *perm, const uint32_t *in, uint32_t *out)
{
#pragma omp parallel for schedule(static) if (np > 65536)
  for (int64_t s = 0; s < np; s++) out[s] = in[perm[s]];
}

void orc_init_particles(const int64_t N[3], const double p0[3], const double d[3],
                        int64_t ppc, uint64_t seed, const double v0[3], const double vT[3],
                        int pert, double amp,
                        double *x, double *y, double *z, double *vx, double *vy, double *vz,
                        uint32_t *pid, double *w)
{
  const int64_t ncell = N[0] * N[1] * N[2];
  const double L[3] = {d[0] * (double)N[0], d[1] * (double)N[1], d[2] * (double)N[2]};
  const double p1[3] = {p0[0] + L[0], p0[1] + L[1], p0[2] + L[2]};
  for (int64_t c = 0; c < ncell; c++) {
    const int64_t i = c % N[0], j = (c / N[0]) % N[1], k = c / (N[0] * N[1]);
    double wc = 1.0;
    if (pert == 2 || pert == 3) {
      double u[4];
      philox_u((uint32_t)c, 0, STREAM_WEIGHT, 0, seed, u);
      wc = (pert == 2) ? 1.0 + amp * (2.0 * u[0] - 1.0)
                       : 1.0 + amp * (sqrt(-2.0 * log(u[0])) * cos(2.0 * M_PI * u[1]));
    }
    for (int64_t s = 0; s < ppc; s++) {
      const int64_t p = c * ppc + s;
      double ua[4], ub[4];
      philox_u((uint32_t)p, 0, STREAM_INIT, 0, seed, ua);
      philox_u((uint32_t)p, 0, STREAM_INIT, 1, seed, ub);
      x[p] = orc_wrap(p0[0] + ((double)i + ua[0]) * d[0], p0[0], p1[0], L[0]);
      y[p] = orc_wrap(p0[1] + ((double)j + ua[1]) * d[1], p0[1], p1[1], L[1]);
      z[p] = orc_wrap(p0[2] + ((double)k + ua[2]) * d[2], p0[2], p1[2], L[2]);
      const double r1 = sqrt(-2.0 * log(ub[0])), a1 = 2.0 * M_PI * ub[1];
      const double r2 = sqrt(-2.0 * log(ub[2])), a2 = 2.0 * M_PI * ub[3];
      vx[p] = v0[0] + vT[0] * (r1 * cos(a1));
      vy[p] = v0[1] + vT[1] * (r1 * sin(a1));
      vz[p] = v0[2] + vT[2] * (r2 * cos(a2));
      pid[p] = (uint32_t)p;
      if (w) {
        if (pert == 1) w[p] = 1.0 + amp * sin(2.0 * M_PI * (x[p] - p0[0]) / L[0]);
        else w[p] = wc;
      }
    }
  }
}

/* ---- the reference's initial densities (particles.c:26-45) -------------------
 * Positions as distributions.c places them; velocities v0 + vT N(0,1) and the
 * weight perturbation from the same Philox blocks as orc_init_particles (the
 * reference draws velocities from gasdev, particles.c:100-126, a stream this
 * build does not reproduce).  Ion p gets id p. */

/* Direction integers, scaled to MAXBIT = 30 bits, of the first three
 * dimensions of the Sobol generator of random.c:234-290: degrees 1, 2, 3
 * (random.c:241), polynomial codes 0, 1, 1 (random.c:242) and the leading odd
 * integers {1}, {1, 1}, {1, 3, 7} (random.c:244).  The later integers follow
 * the recurrence of random.c:260-269 as written there: it XORs in the
 * previous integer V[j-1] once for each set polynomial bit.  For the third
 * dimension this is the recurrence of the primitive polynomial x^3 + x^2 + 1
 * (not the x^3 + x + 1 that code 1 names in the Numerical Recipes table). */
void orc_sobol_dirs(uint32_t V[30][3])
{
  static const int deg[3] = {1, 2, 3};
  static const unsigned code[3] = {0u, 1u, 1u};
  static const uint32_t lead[3][3] = {{1u, 0u, 0u}, {1u, 1u, 0u}, {1u, 3u, 7u}};
  for (int k = 0; k < 3; k++) {
    const int s = deg[k];
    for (int j = 0; j < s; j++) V[j][k] = lead[k][j] << (29 - j);
    for (int j = s; j < 30; j++) {
      uint32_t v = V[j - s][k];
      v ^= v >> s;
      unsigned bits = code[k];
      for (int l = s - 1; l >= 1; l--) {
        if (bits & 1u) v ^= V[j - 1][k];
        bits >>= 1;
      }
      V[j][k] = v;
    }
  }
}

/* The first np points of the sequence in the order successive calls return
 * them (random.c:273-287): call t XORs the direction integer of the lowest
 * zero bit of t into the running integers.  r is [np][3] in [0, 1).  Returns
 * -1 when a call would need a 31st bit (the reference's "MAXBIT too small"). */
int orc_sobol_points(int64_t np, double *r)
{
  uint32_t V[30][3];
  orc_sobol_dirs(V);
  uint32_t ix[3] = {0u, 0u, 0u};
  for (int64_t t = 0; t < np; t++) {
    uint64_t im = (uint64_t)t;
    int j = 0;
    while (im & 1u) { im >>= 1; j++; }
    if (j >= 30) return -1;
    for (int k = 0; k < 3; k++) {
      ix[k] ^= V[j][k];
      r[3 * t + k] = (double)ix[k] * (1.0 / 1073741824.0);
    }
  }
  return 0;
}

/* 53-bit uniform in [0, 1) from two Philox words. */
static double u53(uint32_t hi, uint32_t lo)
{
  return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * (1.0 / 9007199254740992.0);
}

enum { STREAM_PLACE = 4 };

/* Velocity, weight and id of ion p at position (x, y, z). */
static void finish_ion(int64_t p, const int64_t N[3], const double p0[3], const double d[3], uint64_t seed,
                       const double v0[3], const double vT[3], int pert, double amp, double *vx, double *vy,
                       double *vz, uint32_t *pid, double *w, double x, double y, double z)
{
  double ub[4];
  philox_u((uint32_t)p, 0, STREAM_INIT, 1, seed, ub);
  const double r1 = sqrt(-2.0 * log(ub[0])), a1 = 2.0 * M_PI * ub[1];
  const double r2 = sqrt(-2.0 * log(ub[2])), a2 = 2.0 * M_PI * ub[3];
  vx[p] = v0[0] + vT[0] * (r1 * cos(a1));
  vy[p] = v0[1] + vT[1] * (r1 * sin(a1));
  vz[p] = v0[2] + vT[2] * (r2 * cos(a2));
  pid[p] = (uint32_t)p;
  if (!w) return;
  if (pert == 1) {
    w[p] = 1.0 + amp * sin(2.0 * M_PI * (x - p0[0]) / (d[0] * (double)N[0]));
  } else if (pert == 2 || pert == 3) {
    int64_t ci[3];
    const double q[3] = {x, y, z};
    for (int a = 0; a < 3; a++) {
      int64_t i = (int64_t)((q[a] - p0[a]) / d[a]);
      if (i > N[a] - 1) i = N[a] - 1;
      if (i < 0) i = 0;
      ci[a] = i;
    }
    const int64_t c = ci[0] + N[0] * (ci[1] + N[1] * ci[2]);
    double u[4];
    philox_u((uint32_t)c, 0, STREAM_WEIGHT, 0, seed, u);
    w[p] = (pert == 2) ? 1.0 + amp * (2.0 * u[0] - 1.0)
                       : 1.0 + amp * (sqrt(-2.0 * log(u[0])) * cos(2.0 * M_PI * u[1]));
  } else {
    w[p] = 1.0;
  }
}

int64_t orc_density_count(int kind, const int64_t N[3], int64_t Np)
{
  if (kind == ORC_DENSITY_UNIFORM) {
    const int64_t nc = N[0] * N[1] * N[2];
    return (Np / nc) * (N[0] - 1) * (N[1] - 1) * (N[2] - 1);
  }
  return Np;
}

int64_t orc_init_density(int kind, const int64_t N[3], const double p0[3], const double p1[3], const double L[3],
                         const double d[3], int64_t Np, uint64_t seed, const double v0[3], const double vT[3],
                         int pert, double amp, double *x, double *y, double *z, double *vx, double *vy,
                         double *vz, uint32_t *pid, double *w)
{
  const int64_t np = orc_density_count(kind, N, Np);
  if (kind == ORC_DENSITY_FLAT_SOBOL) {
    /* distributions.c:95-141: v = r (L + d) - d/2, pulled in by 0.01 d when it
     * falls outside [p0, p1] */
    double *r = malloc(sizeof(double) * 3 * (size_t)(np > 0 ? np : 1));
    if (orc_sobol_points(np, r)) { free(r); return -1; }
    for (int64_t p = 0; p < np; p++) {
      double q[3];
      for (int a = 0; a < 3; a++) {
        double v = p0[a] + (r[3 * p + a] * (L[a] + d[a]) - 0.5 * d[a]);
        if (v < p0[a]) v = p0[a] + 0.01 * d[a];
        else if (v > p1[a]) v = p1[a] - 0.01 * d[a];
        q[a] = orc_wrap(v, p0[a], p1[a], L[a]);
      }
      x[p] = q[0]; y[p] = q[1]; z[p] = q[2];
      finish_ion(p, N, p0, d, seed, v0, vT, pert, amp, vx, vy, vz, pid, w, q[0], q[1], q[2]);
    }
    free(r);
  } else if (kind == ORC_DENSITY_UNIFORM) {
    /* distributions.c:33-92: Np / nc ions at the centre of every cell except
     * the last one along each axis, cell index k fastest */
    const int64_t mx = N[0] - 1, my = N[1] - 1, mz = N[2] - 1, m = mx * my * mz;
    for (int64_t p = 0; p < np; p++) {
      const int64_t q = p % m;
      const int64_t k = q % mz, j = (q / mz) % my, i = q / (mz * my);
      x[p] = p0[0] + d[0] * ((double)(i + 1) - 0.5);
      y[p] = p0[1] + d[1] * ((double)(j + 1) - 0.5);
      z[p] = p0[2] + d[2] * ((double)(k + 1) - 0.5);
      finish_ion(p, N, p0, d, seed, v0, vT, pert, amp, vx, vy, vz, pid, w, x[p], y[p], z[p]);
    }
  } else if (kind == ORC_DENSITY_SINUSOIDAL) {
    /* distributions.c:144-232 with SinusoidalDistribution (152-162): accept a
     * uniform point where 1 + 0.25 sin(2 pi x) > u * 1.25, the density's
     * maximum; the candidates of ion p come from its own Philox blocks
     * (attempt a: {p, a, PLACE, 0/1}); after 64 rejections the last candidate
     * is kept (probability below 0.4^64) */
    for (int64_t p = 0; p < np; p++) {
      double q[3] = {p0[0], p0[1], p0[2]};
      for (uint32_t a = 0; a < 64; a++) {
        const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        const uint32_t c0[4] = {(uint32_t)p, a, STREAM_PLACE, 0u}, c1[4] = {(uint32_t)p, a, STREAM_PLACE, 1u};
        uint32_t b0[4], b1[4];
        orc_philox4x32_10(c0, key, b0);
        orc_philox4x32_10(c1, key, b1);
        const double u[3] = {u53(b0[0], b0[1]), u53(b0[2], b0[3]), u53(b1[0], b1[1])};
        for (int e = 0; e < 3; e++) q[e] = orc_wrap(p0[e] + u[e] * L[e], p0[e], p1[e], L[e]);
        const double dens = 1.0 + 0.25 * sin(2.0 * M_PI * q[0]);
        if (dens > orc_u24(b1[2]) * 1.25) break;
      }
      x[p] = q[0]; y[p] = q[1]; z[p] = q[2];
      finish_ion(p, N, p0, d, seed, v0, vT, pert, amp, vx, vy, vz, pid, w, q[0], q[1], q[2]);
    }
  } else {
    return -2;
  }
  return np;
}
