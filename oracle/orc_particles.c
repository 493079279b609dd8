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
  double Sn = 0.0;
  stats[0] = stats[1] = stats[2] = 0;
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
      if (r == 1) stats[0]++;
      if (r == 2) stats[1]++;
      stats[2] += capped;
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
    Sn += orc_mean_rel_speed(v, P->vn0, P->vnT);
  }
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
  for (int64_t s = 0; s < np; s++) out[s] = in[perm[s]];
}

void orc_permute_u(int64_t np, const int64_t *perm, const uint32_t *in, uint32_t *out)
{
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
