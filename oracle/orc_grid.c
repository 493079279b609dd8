/*
 * orc_grid.c -- oracle (test infrastructure only): operator, forcing, node
 * field, gather and deposit.  FP64 restatement of the reference formulas;
 * see orc.h for the role of this file.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int64_t wrapi(int64_t i, int64_t n) { i %= n; return i < 0 ? i + n : i; }
static inline int64_t lin(const int64_t N[3], int64_t i, int64_t j, int64_t k)
{
  return wrapi(i, N[0]) + N[0] * (wrapi(j, N[1]) + N[1] * wrapi(k, N[2]));
}

/* Column order of one operator row, reference lhs.c:11-104. */
static const int OFF[19][3] = {
  {+1, 0, 0}, {-1, 0, 0}, {0, +1, 0}, {0, -1, 0}, {0, 0, +1}, {0, 0, -1},
  {+1, +1, 0}, {+1, -1, 0}, {-1, +1, 0}, {-1, -1, 0},
  {+1, 0, +1}, {+1, 0, -1}, {-1, 0, +1}, {-1, 0, -1},
  {0, +1, +1}, {0, +1, -1}, {0, -1, +1}, {0, -1, -1},
  {0, 0, 0}};

void orc_stencil_offsets(int off[19][3]) { memcpy(off, OFF, sizeof(OFF)); }

/* Magnetisation tensor r_ab and scaled coefficients s_ab = r_ab h_ab / detA,
 * reference lhs.c:405-447 (and rhs.c:199-216 for r). */
typedef struct { double sxx, syy, szz, sxy, sxz, syz, syx, szx, szy; } scoef;

static scoef full_scoef(const double d[3], const double K[3])
{
  const double Kx = K[0], Ky = K[1], Kz = K[2];
  const double detA = 1 + Kx * Kx + Ky * Ky + Kz * Kz;
  const double rxx = 1 + Kx * Kx, rxy = Ky * Kx - Kz, rxz = Kz * Kx + Ky;
  const double ryx = Kx * Ky + Kz, ryy = 1 + Ky * Ky, ryz = Kz * Ky - Kx;
  const double rzx = Kx * Kz - Ky, rzy = Ky * Kz + Kx, rzz = 1 + Kz * Kz;
  const double dx = d[0], dy = d[1], dz = d[2];
  const double hxx = 1.0 / (2.0 * dx * dx), hyy = 1.0 / (2.0 * dy * dy), hzz = 1.0 / (2.0 * dz * dz);
  const double hxy = 1.0 / (8.0 * dy * dx), hxz = 1.0 / (8.0 * dz * dx), hyz = 1.0 / (8.0 * dy * dz);
  const double scale = 1.0 / detA;
  scoef s;
  s.sxx = scale * rxx * hxx; s.syy = scale * ryy * hyy; s.szz = scale * rzz * hzz;
  s.sxy = scale * rxy * hxy; s.sxz = scale * rxz * hxz; s.syz = scale * ryz * hyz;
  s.syx = scale * ryx * hxy; s.szx = scale * rzx * hxz; s.szy = scale * rzy * hyz;
  return s;
}

/* Row coefficients from the 7-point density star, lhs.c:479-511. */
static void full_row_from_star(const scoef *s, double nc, double nxp, double nxm,
                               double nyp, double nym, double nzp, double nzm, double v[19])
{
  /* x-y edges */
  v[6] = s->sxy * (nxp + nc) + s->syx * (nyp + nc);
  v[7] = -s->sxy * (nxp + nc) - s->syx * (nc + nym);
  v[8] = -s->sxy * (nc + nxm) - s->syx * (nyp + nc);
  v[9] = s->sxy * (nc + nxm) + s->syx * (nc + nym);
  /* x-z edges */
  v[10] = s->sxz * (nxp + nc) + s->szx * (nzp + nc);
  v[11] = -s->sxz * (nxp + nc) - s->szx * (nc + nzm);
  v[12] = -s->sxz * (nc + nxm) - s->szx * (nzp + nc);
  v[13] = s->sxz * (nc + nxm) + s->szx * (nc + nzm);
  /* y-z edges */
  v[14] = s->syz * (nyp + nc) + s->szy * (nzp + nc);
  v[15] = -s->syz * (nyp + nc) - s->szy * (nc + nzm);
  v[16] = -s->syz * (nc + nym) - s->szy * (nzp + nc);
  v[17] = s->syz * (nc + nym) + s->szy * (nc + nzm);
  /* faces */
  v[0] = s->sxx * (nxp + nc) + s->syx * (nyp - nym) + s->szx * (nzp - nzm);
  v[1] = s->sxx * (nc + nxm) - s->syx * (nyp - nym) - s->szx * (nzp - nzm);
  v[2] = s->syy * (nyp + nc) + s->sxy * (nxp - nxm) + s->szy * (nzp - nzm);
  v[3] = s->syy * (nc + nym) - s->sxy * (nxp - nxm) - s->szy * (nzp - nzm);
  v[4] = s->szz * (nzp + nc) + s->sxz * (nxp - nxm) + s->syz * (nyp - nym);
  v[5] = s->szz * (nc + nzm) - s->sxz * (nxp - nxm) - s->syz * (nyp - nym);
  /* centre */
  v[18] = -(s->sxx * (nxp + 2 * nc + nxm) + s->syy * (nyp + 2 * nc + nym) + s->szz * (nzp + 2 * nc + nzm));
}

void orc_lhs_full_row(const int64_t N[3], const double d[3], const double K[3],
                      const double *n, int64_t c, double coef[19])
{
  const int64_t i = c % N[0], j = (c / N[0]) % N[1], k = c / (N[0] * N[1]);
  const scoef s = full_scoef(d, K);
  full_row_from_star(&s, n[c], n[lin(N, i + 1, j, k)], n[lin(N, i - 1, j, k)],
                     n[lin(N, i, j + 1, k)], n[lin(N, i, j - 1, k)],
                     n[lin(N, i, j, k + 1)], n[lin(N, i, j, k - 1)], coef);
}

/* Laplacian row, lhs.c:205-218: face weights dy*dz/dx etc., centre -sum. */
static void laplacian_row(const double d[3], double v[19])
{
  const double sxx = d[1] * d[2] / d[0], syy = d[0] * d[2] / d[1], szz = d[0] * d[1] / d[2];
  for (int q = 0; q < 19; q++) v[q] = 0.0;
  v[0] = sxx; v[1] = sxx; v[2] = syy; v[3] = syy; v[4] = szz; v[5] = szz;
  v[18] = -(v[0] + v[2] + v[4] + v[1] + v[3] + v[5]);
}

static void row_coeffs(int type, const int64_t N[3], const double d[3], const double K[3],
                       const double *n, int64_t c, double v[19])
{
  if (type == ORC_LHS_FULL) {
    orc_lhs_full_row(N, d, K, n, c, v);
  } else if (type == ORC_LHS_LAPLACIAN) {
    laplacian_row(d, v);
  } else {
    for (int q = 0; q < 19; q++) v[q] = 0.0;
    v[18] = 1.0;
  }
}

void orc_lhs_apply(int type, const int64_t N[3], const double d[3], const double K[3],
                   const double *n, const double *x, double *y)
{
  const int64_t nc = N[0] * N[1] * N[2];
#pragma omp parallel for schedule(static) if (nc > 16384)
  for (int64_t c = 0; c < nc; c++) {
    const int64_t i = c % N[0], j = (c / N[0]) % N[1], k = c / (N[0] * N[1]);
    double v[19], acc = 0.0;
    row_coeffs(type, N, d, K, n, c, v);
    for (int q = 0; q < 19; q++) acc += v[q] * x[lin(N, i + OFF[q][0], j + OFF[q][1], k + OFF[q][2])];
    y[c] = acc;
  }
}

void orc_lhs_apply_transpose(int type, const int64_t N[3], const double d[3], const double K[3],
                             const double *n, const double *x, double *y)
{
  const int64_t nc = N[0] * N[1] * N[2];
  for (int64_t c = 0; c < nc; c++) y[c] = 0.0;
  for (int64_t r = 0; r < nc; r++) {
    const int64_t i = r % N[0], j = (r / N[0]) % N[1], k = r / (N[0] * N[1]);
    double v[19];
    row_coeffs(type, N, d, K, n, r, v);
    for (int q = 0; q < 19; q++) y[lin(N, i + OFF[q][0], j + OFF[q][1], k + OFF[q][2])] += v[q] * x[r];
  }
}

void orc_lhs_dense(int type, const int64_t N[3], const double d[3], const double K[3],
                   const double *n, double *A)
{
  const int64_t nc = N[0] * N[1] * N[2];
  memset(A, 0, sizeof(double) * nc * nc);
  for (int64_t r = 0; r < nc; r++) {
    const int64_t i = r % N[0], j = (r / N[0]) % N[1], k = r / (N[0] * N[1]);
    double v[19];
    row_coeffs(type, N, d, K, n, r, v);
    /* MatSetValuesStencil(INSERT_VALUES): with N >= 3 the 19 columns are distinct. */
    for (int q = 0; q < 19; q++) A[r * nc + lin(N, i + OFF[q][0], j + OFF[q][1], k + OFF[q][2])] = v[q];
  }
}

void orc_remove_mean(int64_t nc, double *v)
{
  double s = 0.0;
  for (int64_t c = 0; c < nc; c++) s += v[c];
  const double m = s / (double)nc;
  for (int64_t c = 0; c < nc; c++) v[c] -= m;
}

/* rhs.c:133-303 with centred differences from calculus.c (dFdx 21-47, d2Fdxx
 * 156-182, d2Fdxy 293-328, d2Fdxz 375-410); d2Fdyz uses the y-z stencil
 * (intended; calculus.c:464 uses the x-z one). */
void orc_rhs(int type, const int64_t N[3], const double d[3], const double K[3],
             const double E0[3], double cth, double cG,
             const double *n, const double *gx, const double *gy, const double *gz,
             double *b)
{
  const int64_t nc = N[0] * N[1] * N[2];
  const double Kx = K[0], Ky = K[1], Kz = K[2];
  const double detA = 1 + Kx * Kx + Ky * Ky + Kz * Kz;
  const double dx = d[0], dy = d[1], dz = d[2];
  if (type == ORC_RHS_CONSTANT) {
    double s = 0.0;
    for (int64_t c = 0; c < nc; c++) s += n[c];
    const double mean = s / (double)nc;
    const double val = mean * (2.0 * dx * dy * dz / detA);
    for (int64_t c = 0; c < nc; c++) b[c] = val;
    return;
  }
  if (type == ORC_RHS_SINUSOIDAL) {
    double s = 0.0;
    for (int64_t c = 0; c < nc; c++) s += n[c];
    const double n0 = s / (double)nc;
    for (int64_t c = 0; c < nc; c++) {
      const int64_t i = c % N[0], j = (c / N[0]) % N[1], k = c / (N[0] * N[1]);
      const double x = ((double)i + 0.5) * dx, y = ((double)j + 0.5) * dy, z = ((double)k + 0.5) * dz;
      const double val = cos(2 * M_PI * x) * cos(2 * M_PI * y) * cos(2 * M_PI * z);
      b[c] = n0 * val * (2.0 * dx * dy * dz / detA);
    }
    orc_remove_mean(nc, b);
    return;
  }
  const double rxx = 1 + Kx * Kx, rxy = Ky * Kx - Kz, rxz = Kz * Kx + Ky;
  const double ryx = Kx * Ky + Kz, ryy = 1 + Ky * Ky, ryz = Kz * Ky - Kx;
  const double rzx = Kx * Kz - Ky, rzy = Ky * Kz + Kx, rzz = 1 + Kz * Kz;
  const double scale = 1.0 / detA;
#pragma omp parallel for schedule(static) if (nc > 16384)
  for (int64_t c = 0; c < nc; c++) {
    const int64_t i = c % N[0], j = (c / N[0]) % N[1], k = c / (N[0] * N[1]);
#define F(a, di, dj, dk) a[lin(N, i + (di), j + (dj), k + (dk))]
    const double dndx = (F(n, 1, 0, 0) - F(n, -1, 0, 0)) / (2.0 * dx);
    const double dndy = (F(n, 0, 1, 0) - F(n, 0, -1, 0)) / (2.0 * dy);
    const double dndz = (F(n, 0, 0, 1) - F(n, 0, 0, -1)) / (2.0 * dz);
    const double dGxdx = (F(gx, 1, 0, 0) - F(gx, -1, 0, 0)) / (2.0 * dx);
    const double dGydy = (F(gy, 0, 1, 0) - F(gy, 0, -1, 0)) / (2.0 * dy);
    const double dGzdz = (F(gz, 0, 0, 1) - F(gz, 0, 0, -1)) / (2.0 * dz);
    const double d2xx = (F(n, 1, 0, 0) - 2.0 * F(n, 0, 0, 0) + F(n, -1, 0, 0)) / (dx * dx);
    const double d2yy = (F(n, 0, 1, 0) - 2.0 * F(n, 0, 0, 0) + F(n, 0, -1, 0)) / (dy * dy);
    const double d2zz = (F(n, 0, 0, 1) - 2.0 * F(n, 0, 0, 0) + F(n, 0, 0, -1)) / (dz * dz);
    const double d2xy = (F(n, 1, 1, 0) - F(n, -1, 1, 0) - F(n, 1, -1, 0) + F(n, -1, -1, 0)) / (4.0 * dx * dy);
    const double d2xz = (F(n, 1, 0, 1) - F(n, -1, 0, 1) - F(n, 1, 0, -1) + F(n, -1, 0, -1)) / (4.0 * dx * dz);
    const double d2yz = (F(n, 0, 1, 1) - F(n, 0, -1, 1) - F(n, 0, 1, -1) + F(n, 0, -1, -1)) / (4.0 * dy * dz);
#undef F
    const double Eterm = (rxx * E0[0] + rxy * E0[1] + rxz * E0[2]) * dndx +
                         (ryx * E0[0] + ryy * E0[1] + ryz * E0[2]) * dndy +
                         (rzx * E0[0] + rzy * E0[1] + rzz * E0[2]) * dndz;
    const double Pterm = cth * (rxx * d2xx + rxy * d2xy + rxz * d2xz +
                                ryx * d2xy + ryy * d2yy + ryz * d2yz +
                                rzx * d2xz + rzy * d2yz + rzz * d2zz);
    const double Gterm = detA * cG * (dGxdx + dGydy + dGzdz);
    b[c] = scale * (Eterm + Pterm + Gterm);
  }
  orc_remove_mean(nc, b);
}

void orc_efield_nodes(const int64_t N[3], const double d[3], const double E0[3],
                      const double *phi, double *ex, double *ey, double *ez)
{
  const int64_t nc = N[0] * N[1] * N[2];
  const double hx = 0.5 / d[0], hy = 0.5 / d[1], hz = 0.5 / d[2];
#pragma omp parallel for schedule(static) if (nc > 16384)
  for (int64_t c = 0; c < nc; c++) {
    const int64_t i = c % N[0], j = (c / N[0]) % N[1], k = c / (N[0] * N[1]);
    ex[c] = E0[0] - (phi[lin(N, i + 1, j, k)] - phi[lin(N, i - 1, j, k)]) * hx;
    ey[c] = E0[1] - (phi[lin(N, i, j + 1, k)] - phi[lin(N, i, j - 1, k)]) * hy;
    ez[c] = E0[2] - (phi[lin(N, i, j, k + 1)] - phi[lin(N, i, j, k - 1)]) * hz;
  }
}

void orc_cell_coord(double x, double p0, double d, int64_t N, int64_t *i, double *f)
{
  const double xi = (x - p0) / d;
  int64_t ii = (int64_t)xi;
  if (ii > N - 1) ii = N - 1;
  if (ii < 0) ii = 0;
  *i = ii;
  *f = xi - (double)ii;
}

/* Corner weights W_q = (wz*wy)*wx for q = dx + 2 dy + 4 dz, the products of
 * particles.c:204-211. */
static inline void corner_weights(double fx, double fy, double fz, double W[8])
{
  const double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy}, wz[2] = {1.0 - fz, fz};
  for (int q = 0; q < 8; q++) W[q] = (wz[q >> 2] * wy[(q >> 1) & 1]) * wx[q & 1];
}

void orc_gather(const int64_t N[3], const double p0[3], const double d[3],
                const double *ex, const double *ey, const double *ez,
                double x, double y, double z, double E[3])
{
  int64_t i, j, k;
  double fx, fy, fz, W[8];
  orc_cell_coord(x, p0[0], d[0], N[0], &i, &fx);
  orc_cell_coord(y, p0[1], d[1], N[1], &j, &fy);
  orc_cell_coord(z, p0[2], d[2], N[2], &k, &fz);
  corner_weights(fx, fy, fz, W);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int q = 0; q < 8; q++) {
    const int64_t c = lin(N, i + (q & 1), j + ((q >> 1) & 1), k + (q >> 2));
    a0 += W[q] * ex[c];
    a1 += W[q] * ey[c];
    a2 += W[q] * ez[c];
  }
  E[0] = a0; E[1] = a1; E[2] = a2;
}

void orc_deposit_sorted(const int64_t N[3], const double p0[3], const double d[3],
                        int64_t np, const double *x, const double *y, const double *z,
                        const double *vx, const double *vy, const double *vz, const double *w,
                        const int64_t *offsets,
                        double *n, double *gx, double *gy, double *gz)
{
  (void)np;
  const int64_t nc = N[0] * N[1] * N[2];
  double *P = (double *)calloc((size_t)nc * 32, sizeof(double));
  /* cells are independent, each summed in its slot order */
#pragma omp parallel for schedule(dynamic, 256) if (nc > 4096)
  for (int64_t c = 0; c < nc; c++) {
    double *Pc = P + 32 * c;
    for (int64_t p = offsets[c]; p < offsets[c + 1]; p++) {
      int64_t i, j, k;
      double fx, fy, fz, W[8];
      orc_cell_coord(x[p], p0[0], d[0], N[0], &i, &fx);
      orc_cell_coord(y[p], p0[1], d[1], N[1], &j, &fy);
      orc_cell_coord(z[p], p0[2], d[2], N[2], &k, &fz);
      corner_weights(fx, fy, fz, W);
      for (int q = 0; q < 8; q++) {
        const double a = w ? w[p] * W[q] : W[q];
        Pc[4 * q + 0] += a;
        Pc[4 * q + 1] += a * vx[p];
        Pc[4 * q + 2] += a * vy[p];
        Pc[4 * q + 3] += a * vz[p];
      }
    }
  }
#pragma omp parallel for schedule(static) if (nc > 16384)
  for (int64_t c = 0; c < nc; c++) {
    const int64_t i = c % N[0], j = (c / N[0]) % N[1], k = c / (N[0] * N[1]);
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < 8; q++) {
      const int64_t s = lin(N, i - (q & 1), j - ((q >> 1) & 1), k - (q >> 2));
      for (int a = 0; a < 4; a++) m[a] += P[32 * s + 4 * q + a];
    }
    n[c] = m[0]; gx[c] = m[1]; gy[c] = m[2]; gz[c] = m[3];
  }
  free(P);
}

void orc_deposit_scatter(const int64_t N[3], const double p0[3], const double d[3],
                         int64_t np, const double *x, const double *y, const double *z,
                         const double *vx, const double *vy, const double *vz, const double *w,
                         double *n, double *gx, double *gy, double *gz)
{
  const int64_t nc = N[0] * N[1] * N[2];
  for (int64_t c = 0; c < nc; c++) n[c] = gx[c] = gy[c] = gz[c] = 0.0;
  for (int64_t p = 0; p < np; p++) {
    int64_t i, j, k;
    double fx, fy, fz, W[8];
    orc_cell_coord(x[p], p0[0], d[0], N[0], &i, &fx);
    orc_cell_coord(y[p], p0[1], d[1], N[1], &j, &fy);
    orc_cell_coord(z[p], p0[2], d[2], N[2], &k, &fz);
    corner_weights(fx, fy, fz, W);
    for (int q = 0; q < 8; q++) {
      const int64_t c = lin(N, i + (q & 1), j + ((q >> 1) & 1), k + (q >> 2));
      const double a = w ? w[p] * W[q] : W[q];
      n[c] += a;
      gx[c] += a * vx[p];
      gy[c] += a * vy[p];
      gz[c] += a * vz[p];
    }
  }
}
